// pfm_context.hip -- the C-ABI of include/pfm.h: system setup, the md() step
// sequence, integrators, Nose-Hoover chains and FP64 reductions on gfx950.
//
// Step order per md_step s (md_simulation.f90:137-231):
//   check_positions, [invert_z]; if s != 0: [nvms exit test], [NHC half],
//   kick, drift; lists; [zero_momentum]; zero_forces; forces; if s != 0:
//   kick, [NHC half], [nvms projection]; [energies].
// All per-atom work runs on the device; the host only sequences launches and
// reads back scalars at energy steps (every step under nvms) and at the end
// of pfm_run.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>

#include "pfm_internal.h"

static int fail(pfm_ctx* h, int code, const std::string& msg) {
  h->err = msg;
  return code;
}

// ------------------------------------------------------------------ timing

static hipEvent_t ev_get(pfm_ctx* h) {
  if (h->evnext >= h->evpool.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    h->evpool.push_back(e);
  }
  return h->evpool[h->evnext++];
}
// the launch records both events (hipExtLaunchKernelGGL)
void tim_begin(pfm_ctx* h, const char* name, hipEvent_t* a, hipEvent_t* b) {
  (void)name;
  *a = ev_get(h);
  *b = ev_get(h);
}
void tim_end(pfm_ctx* h, hipEvent_t a, hipEvent_t b, const char* name) {
  h->timers.push_back(KTimer{name, a, b});
}

// ------------------------------------------------------------ device kernels


struct NhcDev {
  double x[16], v[16], q[16];
  double s, e, kt, L, T;
  int M;
  uint32_t bit;
};

// first half of a step for every atom (kick_drift_atom)
__global__ void k_kick_drift(int N, double4* __restrict__ pos, double4* __restrict__ vel,
                             double4* __restrict__ frc, const uint32_t* __restrict__ mask, IntP P, DevErr* err,
                             const int* __restrict__ orig, int step, uint4* __restrict__ qpos, double3 qs,
                             const int* __restrict__ rep, const RepDev* __restrict__ rdev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  kick_drift_atom(i, pos[i], mask[i], P, rep, rdev, qs, pos, vel, frc, qpos, false, err, orig, step);
}

// fixed-point copy of the positions (mixed path), refreshed at the start of
// every run so positions set from the host are picked up
__global__ void k_quantize(int N, const double4* __restrict__ pos, uint4* __restrict__ qpos, double3 qs,
                           const int* __restrict__ rep, const RepDev* __restrict__ rdev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) qpos[i] = quantize(pos[i], rep ? rdev[rep[i]].qs : qs, 0u);
}

__global__ void k_check(int N, const double4* __restrict__ pos, double bx, double by, double bz, DevErr* err,
                        const int* __restrict__ orig, int step, int invert, double4* __restrict__ vel, double zlo,
                        double zhi, const int* __restrict__ rep, const RepDev* __restrict__ rdev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (rep) {
    const RepDev& R = rdev[rep[i]];
    bx = R.L.x; by = R.L.y; bz = R.L.z; zlo = R.zlo; zhi = R.zhi;
  }
  double4 p = pos[i];
  if (!(in_cell(p.x, bx) && in_cell(p.y, by) && in_cell(p.z, bz))) raise_out_of_cell(err, orig[i] + 1, step, p);
  if (invert) {
    double4 v = vel[i];
    invert_v(v, p.z, zlo, zhi);
    vel[i] = v;
  }
}

#define RED_THREADS 256
#define RED_MAXQ 8

struct MaskQ {
  int nq;
  uint32_t bit[RED_MAXQ];
};

// compensated sum of one (s, c) pair per thread over the block, in a fixed
// order: a butterfly over each wavefront's 64 lanes (cross-lane shuffles, no
// barrier), then the four wavefront sums in LDS in wavefront order.  Every
// thread returns the block's pair.  sh holds at least 8 doubles (callers
// pass 2 * RED_THREADS).
__device__ __forceinline__ CSum block_csum(CSum v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v.add(CSum{__shfl_xor(v.s, o), __shfl_xor(v.c, o)});
  const int w = threadIdx.x >> 6, nw = RED_THREADS >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w] = v.s;
    sh[nw + w] = v.c;
  }
  __syncthreads();
  CSum r{sh[0], sh[nw]};
  for (int k = 1; k < nw; ++k) r.add(CSum{sh[k], sh[nw + k]});
  __syncthreads();
  return r;
}

// Per-block partials of compensated sums: row q of a grid of nb blocks keeps
// the s parts at part[2q nb + b] and the c parts at part[(2q+1) nb + b].
__device__ __forceinline__ void put_part(double* part, int q, int nb, int b, const CSum& v) {
  part[(size_t)(2 * q) * nb + b] = v.s;
  part[(size_t)(2 * q + 1) * nb + b] = v.c;
}
// the compensated sum of row q's nb partials, in a fixed order
__device__ __forceinline__ CSum reduce_part(const double* part, int q, int nb, double* sh) {
  CSum a{0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += RED_THREADS)
    a.add(CSum{part[(size_t)(2 * q) * nb + b], part[(size_t)(2 * q + 1) * nb + b]});
  return block_csum(a, sh);
}

// second half: kick xyz/z, then for nvms the projections
// molecular_static_xyz/1D_velocities (md_integrators.f90:99-145)
__global__ void k_kick2(int N, double4* __restrict__ vel, const double4* __restrict__ frc,
                        const uint32_t* __restrict__ mask, IntP P, int nvms) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  uint32_t m = mask[i];
  double4 v = vel[i], f = frc[i];  // loaded beside the mask, not after it
  bool bx = (m & P.bxyz) != 0, bz = (m & P.bz) != 0;
  if (!(bx || bz)) return;
  double k = 1.0 / v.w / PFM_MASS_COEF;
  if (bx) { v.x = v.x + f.x * k * P.ts1; v.y = v.y + f.y * k * P.ts1; v.z = v.z + f.z * k * P.ts1; }
  if (bz) v.z = v.z + f.z * k * P.ts1;
  if (nvms) {
    if (bx) {
      double fv = f.x * v.x + f.y * v.y + f.z * v.z, ff = f.x * f.x + f.y * f.y + f.z * f.z;
      if (fv > 0.0 && ff > 1e-12) { double c = fv / ff; v.x = c * f.x; v.y = c * f.y; v.z = c * f.z; }
      else { v.x = 0; v.y = 0; v.z = 0; }
    }
    if (bz) {
      double fv = f.x * v.x + f.y * v.y + f.z * v.z;
      if (!(fv > 0.0)) { v.x = 0; v.y = 0; v.z = 0; }
    }
  }
  vel[i] = v;
}



__global__ void k_scale(int N, double4* __restrict__ vel, const uint32_t* __restrict__ mask, IntP P) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  uint32_t m = mask[i];
  for (int c = 0; c < P.nchain; ++c)
    if (m & P.chain_bit[c]) {
      double4 v = vel[i];
      double s = P.chain_s[c];
      v.x *= s; v.y *= s; v.z *= s;
      vel[i] = v;
    }
}

__global__ void k_zero_masked(int N, double4* __restrict__ frc, const uint32_t* __restrict__ mask, uint32_t bit) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N && (mask[i] & bit)) frc[i] = make_double4(0, 0, 0, 0);
}


// per-block partial kinetic energies m v.v/2*mass_coef of nq groups
// (calculate_kinetic_energy, md_general.f90:161-182)
__global__ void __launch_bounds__(RED_THREADS) k_ke_part(int N, const double4* __restrict__ vel,
                                                         const uint32_t* __restrict__ mask, MaskQ Q,
                                                         double* __restrict__ part) {
  __shared__ double sh[2 * RED_THREADS];
  CSum acc[RED_MAXQ];
  for (int q = 0; q < Q.nq; ++q) acc[q] = CSum{0.0, 0.0};
  for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < N; i += gridDim.x * RED_THREADS) {
    double4 v = vel[i];
    uint32_t m = mask[i];
    double ke = v.w * (v.x * v.x + v.y * v.y + v.z * v.z) / 2 * PFM_MASS_COEF;
    for (int q = 0; q < Q.nq; ++q)
      if (m & Q.bit[q]) acc[q].add(ke);
  }
  for (int q = 0; q < Q.nq; ++q) {
    CSum s = block_csum(acc[q], sh);
    if (threadIdx.x == 0) put_part(part, q, gridDim.x, blockIdx.x, s);
  }
}

// per-block partial sums of m*v (3) and m over one group
__global__ void __launch_bounds__(RED_THREADS) k_mom_part(int N, const double4* __restrict__ vel,
                                                          const uint32_t* __restrict__ mask, uint32_t bit,
                                                          double* __restrict__ part) {
  __shared__ double sh[2 * RED_THREADS];
  CSum a[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < N; i += gridDim.x * RED_THREADS) {
    if (mask[i] & bit) {
      double4 v = vel[i];
      a[0].add(v.w * v.x); a[1].add(v.w * v.y); a[2].add(v.w * v.z); a[3].add(v.w);
    }
  }
  for (int q = 0; q < 4; ++q) {
    CSum s = block_csum(a[q], sh);
    if (threadIdx.x == 0) put_part(part, q, gridDim.x, blockIdx.x, s);
  }
}

// reduce nq rows of nb partials each (fixed order) into out[q]
__global__ void __launch_bounds__(RED_THREADS) k_reduce_rows(int nq, int nb, const double* __restrict__ part,
                                                             double* __restrict__ out) {
  __shared__ double sh[2 * RED_THREADS];
  for (int q = 0; q < nq; ++q) {
    CSum s = reduce_part(part, q, nb, sh);
    if (threadIdx.x == 0) out[q] = s.value();
  }
}

// zero_momentum: v -= sum(m v)/sum(m) over the group (md_general.f90:236-253)
__global__ void k_sub_mcv(int N, double4* __restrict__ vel, const uint32_t* __restrict__ mask, uint32_t bit,
                          const double* __restrict__ sums) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !(mask[i] & bit)) return;
  double tm = sums[3];
  double4 v = vel[i];
  v.x -= sums[0] / tm; v.y -= sums[1] / tm; v.z -= sums[2] / tm;
  vel[i] = v;
}

// ------------------------------------------------------------- records

// Records count themselves on the device: ctr = {energy records, frames};
// record k goes to slot k % capacity of its ring.  The last kernel of a
// record advances the counts.
__global__ void k_rec_bump(int* ctr, int e, int x) {
  if (threadIdx.x != 0) return;
  if (e) ctr[0] = ctr[0] + 1;
  if (x) ctr[1] = ctr[1] + 1;
}

struct RecPartArgs {
  int ni, N;
  const double* e_row[PFM_MAXI];  // per-row energies of every interaction
  int e_rows[PFM_MAXI];
  const double4 *pos, *vel, *frc;
  const uint32_t* mask;
  uint32_t ke_bit, all_bit;  // all_moving, all_atoms
  double* epart;             // [ni + 1] compensated partial rows: interaction energies, kinetic energy
  double* opart;             // [10] compensated partial rows of the out block and [nb] largest
                             // v.v (at 20 nb), or null for the energies alone
};

// Per-block partials of one energy record.  The energy rows and the kinetic
// energy are summed in the order of k_sum_part and k_ke_part (the values of
// pfm_energies); the out block of md(): rows 0-2 the force sum, 3-5 sum m x,
// 6-8 sum m v, 9 sum m over the all_atoms group (calculate_force_sum,
// calculate_mass_center[_velocity], md_general.f90:184-234,277-298), row 10
// the largest v.v over every atom (find_max_velocity, md_general.f90:366-380).
__global__ void __launch_bounds__(RED_THREADS) k_rec_part(RecPartArgs A) {
  __shared__ double sh[2 * RED_THREADS];
  const int nb = gridDim.x, b = blockIdx.x, t = threadIdx.x;
  for (int k = 0; k < A.ni; ++k) {
    CSum a{0.0, 0.0};
    for (int i = b * RED_THREADS + t; i < A.e_rows[k]; i += nb * RED_THREADS) a.add(A.e_row[k][i]);
    CSum s = block_csum(a, sh);
    if (t == 0) put_part(A.epart, k, nb, b, s);
  }
  CSum ke{0.0, 0.0};
  CSum a[10];
  for (int q = 0; q < 10; ++q) a[q] = CSum{0.0, 0.0};
  double vmax = -1.0;
  for (int i = b * RED_THREADS + t; i < A.N; i += nb * RED_THREADS) {
    double4 v = A.vel[i];
    uint32_t m = A.mask[i];
    double v2 = v.x * v.x + v.y * v.y + v.z * v.z;
    vmax = v2 > vmax ? v2 : vmax;
    if (m & A.ke_bit) ke.add(v.w * (v.x * v.x + v.y * v.y + v.z * v.z) / 2 * PFM_MASS_COEF);
    if (A.opart && (m & A.all_bit)) {
      double4 p = A.pos[i], f = A.frc[i];
      a[0].add(f.x); a[1].add(f.y); a[2].add(f.z);
      a[3].add(v.w * p.x); a[4].add(v.w * p.y); a[5].add(v.w * p.z);
      a[6].add(v.w * v.x); a[7].add(v.w * v.y); a[8].add(v.w * v.z);
      a[9].add(v.w);
    }
  }
  CSum s = block_csum(ke, sh);
  if (t == 0) put_part(A.epart, A.ni, nb, b, s);
  if (!A.opart) return;
  for (int q = 0; q < 10; ++q) {
    s = block_csum(a[q], sh);
    if (t == 0) put_part(A.opart, q, nb, b, s);
  }
  sh[t] = vmax;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] = sh[t + o] > sh[t] ? sh[t + o] : sh[t];
    __syncthreads();
  }
  if (t == 0) A.opart[20 * nb + b] = sh[0];
}

// One energy record into its slot: the ne energy rows reduced as
// k_reduce_rows does, the out-block rows, and a copy of the chain states;
// then the record counts advance (the frame's too when the step has one).
__global__ void __launch_bounds__(RED_THREADS) k_rec_energy(int ne, int nb, const double* __restrict__ epart,
                                                            const double* __restrict__ opart,
                                                            const double* __restrict__ chains, int nchain_dbl,
                                                            double* __restrict__ slots, int width, int ecap,
                                                            int* __restrict__ ctr, int frame) {
  __shared__ double sh[2 * RED_THREADS];
  const int t = threadIdx.x;
  const int n_e = ctr[0];
  double* out = slots + (size_t)(n_e % ecap) * width;
  for (int q = 0; q < ne + 10; ++q) {
    CSum s = q < ne ? reduce_part(epart, q, nb, sh) : reduce_part(opart, q - ne, nb, sh);
    if (t == 0) out[q] = s.value();
  }
  double m = -1.0;
  for (int b = t; b < nb; b += RED_THREADS) m = opart[20 * nb + b] > m ? opart[20 * nb + b] : m;
  sh[t] = m;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] = sh[t + o] > sh[t] ? sh[t + o] : sh[t];
    __syncthreads();
  }
  if (t == 0) out[ne + 10] = sh[0];
  for (int j = t; j < nchain_dbl; j += RED_THREADS) out[ne + 11 + j] = chains[j];
  __syncthreads();
  if (t == 0) {
    ctr[0] = n_e + 1;
    if (frame) ctr[1] = ctr[1] + 1;
  }
}

// One frame in original atom order into device staging: positions then
// velocities, double[n][3] each
__global__ void k_rec_frame(int N, const double4* __restrict__ pos, const double4* __restrict__ vel,
                            const int* __restrict__ orig, double* __restrict__ stage) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t a = 3 * (size_t)orig[i];
  double4 p = pos[i], v = vel[i];
  stage[a] = p.x; stage[a + 1] = p.y; stage[a + 2] = p.z;
  stage[3 * (size_t)N + a] = v.x; stage[3 * (size_t)N + a + 1] = v.y; stage[3 * (size_t)N + a + 2] = v.z;
}

// the staged frame into its pinned host slot, in order, 16 bytes per store
__global__ void k_rec_frame_out(size_t n2, const double2* __restrict__ stage, double* __restrict__ slots, int xcap,
                                const int* __restrict__ ctr) {
  double2* out = (double2*)(slots + (size_t)(ctr[1] % xcap) * 2 * n2);
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n2; k += (size_t)gridDim.x * blockDim.x)
    out[k] = stage[k];
}

// Batches: replica r owns internal atoms [row0[r], row0[r+1]) (slot 0 of the
// row-offset table) and, in group g, rows [row0[g*(nrep+1)+r], ...+1); one
// block per replica.

// zero_momentum (md_general.f90:96-112) of every replica on its own
__global__ void __launch_bounds__(RED_THREADS) k_rep_zero_mom(int nrep, const int* __restrict__ row0,
                                                              double4* __restrict__ vel,
                                                              const uint32_t* __restrict__ mask, uint32_t bit) {
  __shared__ double sh[2 * RED_THREADS];
  const int r = blockIdx.x;
  const int a0 = row0[r], a1 = row0[r + 1];
  CSum c[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int i = a0 + threadIdx.x; i < a1; i += RED_THREADS) {
    if (mask[i] & bit) {
      double4 v = vel[i];
      c[0].add(v.w * v.x); c[1].add(v.w * v.y); c[2].add(v.w * v.z); c[3].add(v.w);
    }
  }
  double a[4];
  for (int q = 0; q < 4; ++q) a[q] = block_csum(c[q], sh).value();
  if (a[3] == 0.0) return;
  for (int i = a0 + threadIdx.x; i < a1; i += RED_THREADS) {
    if (mask[i] & bit) {
      double4 v = vel[i];
      v.x -= a[0] / a[3]; v.y -= a[1] / a[3]; v.z -= a[2] / a[3];
      vel[i] = v;
    }
  }
}

struct RepEnergyArgs {
  int ni, nrep;
  const double* e_row[PFM_MAXI];
  int g1[PFM_MAXI];
  const int* row0;
  const double4* vel;
  const uint32_t* mask;
  uint32_t bit;
  double* out;  // [nrep][ni + 1]: interaction energies, kinetic energy
};

__global__ void __launch_bounds__(RED_THREADS) k_rep_energy(RepEnergyArgs A) {
  __shared__ double sh[2 * RED_THREADS];
  const int r = blockIdx.x, w = A.nrep + 1;
  for (int k = 0; k < A.ni; ++k) {
    const int r0 = A.row0[A.g1[k] * w + r], r1 = A.row0[A.g1[k] * w + r + 1];
    CSum e{0.0, 0.0};
    for (int q = r0 + threadIdx.x; q < r1; q += RED_THREADS) e.add(A.e_row[k][q]);
    e = block_csum(e, sh);
    if (threadIdx.x == 0) A.out[r * (A.ni + 1) + k] = e.value();
  }
  CSum ke{0.0, 0.0};
  for (int i = A.row0[r] + threadIdx.x; i < A.row0[r + 1]; i += RED_THREADS) {
    if (A.mask[i] & A.bit) {
      double4 v = A.vel[i];
      ke.add(v.w * (v.x * v.x + v.y * v.y + v.z * v.z) / 2 * PFM_MASS_COEF);
    }
  }
  ke = block_csum(ke, sh);
  if (threadIdx.x == 0) A.out[r * (A.ni + 1) + A.ni] = ke.value();
}

// integrate_nose_hoover_chain (md_integrators.f90:200-245) on one chain with
// the chain group's kinetic energy ke; returns the velocity scale s
__device__ double nhc_update(NhcDev& h, double ke, double ts1, double ts2, double ts3) {
  double kt = h.kt;
  double kedif = 2.0 * ke - 3.0 * h.L * kt;
  int M = h.M;
  double b = 1.0;
  double* v = h.v;
  const double* q = h.q;
  if (M == 1) {
    v[0] = v[0] + kedif / q[0] * ts2;
  } else {
    v[M - 1] = v[M - 1] + (q[M - 2] * v[M - 2] * v[M - 2] - kt) / q[M - 1] * ts2;
    for (int i = M - 2; i >= 1; --i) {
      b = exp(-v[i + 1] * ts3);
      v[i] = v[i] * (b * b) + (q[i - 1] * v[i - 1] * v[i - 1] - kt) / q[i] * ts2 * b;
    }
    b = exp(-v[1] * ts3);
    v[0] = v[0] * (b * b) + kedif / q[0] * ts2 * b;
  }
  h.s = exp(-v[0] * ts1);
  kedif = 2.0 * ke * (h.s * h.s) - 3.0 * h.L * kt;
  for (int i = 0; i < M; ++i) h.x[i] = h.x[i] + v[i] * ts1;
  if (M == 1) {
    v[0] = v[0] + kedif / q[0] * ts2;
  } else {
    v[0] = v[0] * (b * b) + kedif / q[0] * ts2 * b;
    for (int i = 1; i <= M - 2; ++i) {
      b = exp(-v[i + 1] * ts3);
      v[i] = v[i] * (b * b) + (q[i - 1] * v[i - 1] * v[i - 1] - kt) / q[i] * ts2 * b;
    }
    v[M - 1] = v[M - 1] + (q[M - 2] * v[M - 2] * v[M - 2] - kt) / q[M - 1] * ts2;
  }
  return h.s;
}

// Reduce each chain's kinetic-energy partials (fixed order) and apply one NHC
// half step (twice = 0), or the second half of this step followed by the first
// half of the next (twice = 1).  Nothing changes the velocities between those
// two halves except invert_z_velocities, which keeps the kinetic energy, so
// the second one uses s^2 * ke; s_out receives the product of the scales,
// applied to the velocities by the next kick.  One block.
struct ChainArgs {
  NhcDev* C;
  double* s_out;
  double ts1, ts2, ts3;
  int twice;
};

__global__ void __launch_bounds__(RED_THREADS) k_nhc_chain(int nchain, ChainArgs A, const double* __restrict__ part,
                                                           int nb) {
  __shared__ double sh[2 * RED_THREADS];
  // the chain's serial update reads and writes its state many times over:
  // work on a copy in LDS, loaded and stored by the whole block at once
  __shared__ NhcDev st;
  constexpr int W = sizeof(NhcDev) / sizeof(double);
  for (int c = 0; c < nchain; ++c) {
    for (int j = threadIdx.x; j < W; j += RED_THREADS) ((double*)&st)[j] = ((const double*)&A.C[c])[j];
    const double ke = reduce_part(part, c, nb, sh).value();  // ends with a barrier: the copy is complete
    if (threadIdx.x == 0) {
      double s = nhc_update(st, ke, A.ts1, A.ts2, A.ts3);
      if (A.twice) s *= nhc_update(st, s * s * ke, A.ts1, A.ts2, A.ts3);
      A.s_out[c] = s;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < W; j += RED_THREADS) ((double*)&A.C[c])[j] = ((const double*)&st)[j];
    __syncthreads();
  }
}



// second half with the kinetic energies of the chain groups after the kick
// (the input of the second NHC half step), per block, grid-stride
__global__ void __launch_bounds__(RED_THREADS) k_kick2_ke(int N, double4* __restrict__ vel,
                                                          const double4* __restrict__ frc,
                                                          const uint32_t* __restrict__ mask, IntP P, MaskQ Q,
                                                          double* __restrict__ part) {
  __shared__ double sh[2 * RED_THREADS];
  CSum acc[RED_MAXQ];
  for (int q = 0; q < Q.nq; ++q) acc[q] = CSum{0.0, 0.0};
  for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < N; i += gridDim.x * RED_THREADS) {
    uint32_t m = mask[i];
    bool bx = (m & P.bxyz) != 0, bz = (m & P.bz) != 0;
    double4 v = vel[i];
    const double4 f = frc[i];  // not waiting on the mask
    if (bx || bz) {
      double k = 1.0 / v.w / PFM_MASS_COEF;
      if (bx) { v.x = v.x + f.x * k * P.ts1; v.y = v.y + f.y * k * P.ts1; v.z = v.z + f.z * k * P.ts1; }
      if (bz) v.z = v.z + f.z * k * P.ts1;
      vel[i] = v;
    }
    double ke = v.w * (v.x * v.x + v.y * v.y + v.z * v.z) / 2 * PFM_MASS_COEF;
    for (int q = 0; q < Q.nq; ++q)
      if (m & Q.bit[q]) acc[q].add(ke);
  }
  for (int q = 0; q < Q.nq; ++q) {
    CSum s = block_csum(acc[q], sh);
    if (threadIdx.x == 0) put_part(part, q, gridDim.x, blockIdx.x, s);
  }
}

// ------------------------------------------------------------- host helpers

static uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) {
    r |= (uint64_t)((x >> b) & 1) << (3 * b);
    r |= (uint64_t)((y >> b) & 1) << (3 * b + 1);
    r |= (uint64_t)((z >> b) & 1) << (3 * b + 2);
  }
  return r;
}

template <class T>
static int dalloc(pfm_ctx* h, T** p, size_t n) {
  HIPCHK(h, hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
  HIPCHK(h, hipMemsetAsync(*p, 0, sizeof(T) * (n ? n : 1), h->stream));
  return PFM_OK;
}

static bool is_cos(int k) { return k == PFM_LJC || k == PFM_MORSEC; }

// ---------------------------------------------------------- group growth

// active atoms of a growing group: the first n of its members (in a batch,
// the first rn[replica] of each replica's members)
__global__ void k_group_count(int N, const int* __restrict__ orig, const int* __restrict__ rank, uint32_t bit, int n,
                              const int* __restrict__ rn, const int* __restrict__ rep, uint32_t* __restrict__ mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int r = rank[orig[i]];
  if (r < 0) return;
  const int lim = rn ? rn[rep[i]] : n;
  uint32_t m = mask[i];
  mask[i] = r < lim ? (m | bit) : (m & ~bit);
}

// change_particle_group_N (md_general.f90:82-94) on host counts
static void change_counts(int* n, const int* size, const std::vector<pfm_ctx::GroupChange>& ch, int step) {
  for (const auto& c : ch) {
    int& g = n[c.to];
    if (step <= c.ts1) {
      g = n[c.from];
      if (step == c.ts1) g = g + 1;
    } else if (step < c.ts2 && (step - c.ts1) % c.frec == 0) {
      g = g + 1;
    }
    if (g > size[c.to]) g = size[c.to];
  }
}

// whether the changes at step s alter any count, starting from the current ones
static bool counts_change_at(const pfm_ctx* h, int s) {
  if (h->changes.empty()) return false;
  if (h->nrep > 1) {
    for (int r = 0; r < h->nrep; ++r) {
      int n[PFM_MAXG + 1], n0[PFM_MAXG + 1], size[PFM_MAXG + 1];
      for (int g = 0; g <= PFM_MAXG; ++g) {
        n0[g] = n[g] = h->grp.rn[g].empty() ? 0 : h->grp.rn[g][r];
        size[g] = h->grp.rsize[g].empty() ? 0 : h->grp.rsize[g][r];
      }
      change_counts(n, size, h->changes, s);
      if (!std::equal(n, n + PFM_MAXG + 1, n0)) return true;
    }
    return false;
  }
  int n[PFM_MAXG + 1];
  std::copy(h->grp.n, h->grp.n + PFM_MAXG + 1, n);
  change_counts(n, h->grp.size, h->changes, s);
  return !std::equal(n, n + PFM_MAXG + 1, h->grp.n);
}

// the changes of step s: host counts, then the active bits of the groups that
// moved.  A batch applies the rule to every replica's own counts (each
// replica is its own md() run); the group's count is their sum.
static void apply_changes(pfm_ctx* h, int s) {
  int old[PFM_MAXG + 1];
  std::copy(h->grp.n, h->grp.n + PFM_MAXG + 1, old);
  // a batch's group moved when any replica's count moved (the per-replica
  // counts can change while their sum stays the same)
  bool moved[PFM_MAXG + 1] = {false};
  if (h->nrep > 1) {
    for (int r = 0; r < h->nrep; ++r) {
      int n[PFM_MAXG + 1], size[PFM_MAXG + 1];
      for (int g = 0; g <= PFM_MAXG; ++g) {
        n[g] = h->grp.rn[g].empty() ? 0 : h->grp.rn[g][r];
        size[g] = h->grp.rsize[g].empty() ? 0 : h->grp.rsize[g][r];
      }
      change_counts(n, size, h->changes, s);
      for (int g = 0; g <= PFM_MAXG; ++g)
        if (!h->grp.rn[g].empty()) {
          moved[g] = moved[g] || h->grp.rn[g][r] != n[g];
          h->grp.rn[g][r] = n[g];
        }
    }
    for (int g = 1; g <= PFM_MAXG; ++g) {
      if (h->grp.rn[g].empty()) continue;
      int tot = 0;
      for (int r = 0; r < h->nrep; ++r) tot += h->grp.rn[g][r];
      h->grp.n[g] = tot;
    }
  } else {
    change_counts(h->grp.n, h->grp.size, h->changes, s);
  }
  for (int g = 1; g <= PFM_MAXG; ++g)
    if ((h->grp.n[g] != old[g] || moved[g]) && h->grp.rank[g]) {
      if (h->grp.d_rn[g])
        (void)hipMemcpyAsync(h->grp.d_rn[g], h->grp.rn[g].data(), sizeof(int) * h->nrep, hipMemcpyHostToDevice,
                             h->stream);
      PFM_LAUNCH(h, "group_count", k_group_count, (h->N + 255) / 256, 256, h->N, h->orig, h->grp.rank[g],
                 1u << (g - 1), h->grp.n[g], h->grp.d_rn[g], h->rep, h->mask);
    }
}

// Groups that grow keep their members under a spare group bit for the list
// rows.  Growth is taken for roles, chain groups and the groups of every
// interaction whose lists come from the generic search (lj, lj1g, rjl, tb,
// rebosc, ljc/morsec): rows of members that have not joined stay empty and the
// candidates are the active atoms, as find_neighbours's loops to group%N
// (md_neighbours.f90:67,70).  A growing graphene group of ljc/morsec with a
// tb list stops at its first step as the reference's does: the members that
// have not joined have no nearest neighbours (graphenenorm.f90:30).
static int growth_setup(pfm_ctx* h) {
  if (h->changes.empty()) return PFM_OK;
  bool grows[PFM_MAXG + 1] = {false};
  for (const auto& c : h->changes) grows[c.to] = true;
  for (auto& it : h->inter)
    for (int k = 0; k < it.nl_n; ++k)
      if ((grows[it.nl[k].g1] || grows[it.nl[k].g2]) && it.kind == PFM_SUR)
        return fail(h, PFM_E_INVALID, "error: a changing group in the lists of a surrogate");
  int spare = PFM_MAXG;
  auto taken = [&](int g) {
    if (!h->grp.idx[g].empty() || h->grp.used[g] || h->grp.shadow[g]) return true;
    for (int k = 0; k < 4; ++k)
      if (h->roles[k] == g) return true;
    for (auto& c : h->nhc)
      if (c.group == g) return true;
    for (int q = 1; q <= PFM_MAXG; ++q)
      if (h->grp.shadow[q] == g) return true;
    return false;
  };
  for (int g = 1; g <= PFM_MAXG; ++g) {
    if (!grows[g]) continue;
    while (spare >= 1 && taken(spare)) --spare;
    if (spare < 1) return fail(h, PFM_E_INVALID, "error: no spare group number for a changing group");
    h->grp.shadow[g] = spare--;
    h->grp.used[g] = true;
  }
  return PFM_OK;
}

// Allocate lists, cells and reductions once all interactions are known.
static int finalize(pfm_ctx* h) {
  if (h->sorted_once) return PFM_OK;
  int rc;
  const int nrep = h->nrep;
  if ((rc = growth_setup(h))) return rc;
  // replica of every original atom; atoms of every group per replica
  std::vector<int> rep_of(h->N);
  for (int r = 0; r < nrep; ++r)
    for (int k = 0; k < h->rep_n[r]; ++k) rep_of[h->rep_start[r] + k] = r;
  std::vector<std::vector<int>> gcnt(PFM_MAXG + 1, std::vector<int>(nrep, 0));
  for (int g = 1; g <= PFM_MAXG; ++g)
    for (int a : h->grp.idx[g]) ++gcnt[g][rep_of[a]];
  if (nrep > 1 && !h->changes.empty())
    for (int g = 1; g <= PFM_MAXG; ++g) {
      h->grp.rn[g] = gcnt[g];  // every member active until the first step applies the changes
      h->grp.rsize[g] = gcnt[g];
    }
  // cell size of every replica: the largest list r_cut among the lists whose
  // groups have atoms in it
  std::vector<double> rmax(nrep, 0.0);
  for (auto& it : h->inter) {
    for (int k = 0; k < it.nl_n; ++k) {
      ListDev& L = it.nl[k];
      if (k == 2 && it.nn_from_tb >= 0) continue;
      double rc_eff = (k == 1 && has_converse(it.kind)) ? it.nl[0].rcut : L.rcut;
      for (int r = 0; r < nrep; ++r)
        if (gcnt[L.g1][r] > 0 || gcnt[L.g2][r] > 0) rmax[r] = std::max(rmax[r], rc_eff);
    }
  }
  // grids, Morton ranks offset by each replica's first cell
  h->hrep.assign(nrep, RepDev());
  std::vector<int> rank_all, xyz_all, crep_all;
  h->ncell = 0;
  h->any_tile = h->any_rows = false;
  for (int r = 0; r < nrep; ++r) {
    RepDev& R = h->hrep[r];
    memset(&R, 0, sizeof R);
    const double* bx = &h->rep_box[3 * r];
    double rm = rmax[r] > 0.0 ? rmax[r] : 1.0;
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = std::max(1, std::min((int)std::floor(bx[d] / rm), 1024));
    // at most 2^21 cells: larger cells only add candidates
    while ((long)nc[0] * nc[1] * nc[2] > (1L << 21)) {
      int d = (nc[0] >= nc[1] && nc[0] >= nc[2]) ? 0 : (nc[1] >= nc[2] ? 1 : 2);
      nc[d] = std::max(1, nc[d] / 2);
    }
    R.L = make_double3(bx[0], bx[1], bx[2]);
    R.iL = make_double3(1.0 / bx[0], 1.0 / bx[1], 1.0 / bx[2]);
    R.hb = make_double3(0.5 * bx[0], 0.5 * bx[1], 0.5 * bx[2]);
    R.qs = make_double3(4294967296.0 / bx[0], 4294967296.0 / bx[1], 4294967296.0 / bx[2]);
    R.hq = make_float3((float)(bx[0] / 4294967296.0), (float)(bx[1] / 4294967296.0), (float)(bx[2] / 4294967296.0));
    R.hiso = (R.hq.x == R.hq.y && R.hq.y == R.hq.z) ? R.hq.x : 0.0f;
    double cl[3];
    for (int d = 0; d < 3; ++d) {
      R.nc[d] = nc[d];
      cl[d] = bx[d] / nc[d];
    }
    R.clen = make_double3(cl[0], cl[1], cl[2]);
    R.inv_clen = make_double3(1.0 / cl[0], 1.0 / cl[1], 1.0 / cl[2]);
    R.cell_base = h->ncell;
    R.tile = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
    R.zlo = 0.8 * bx[2];
    R.zhi = 0.9 * bx[2];
    (R.tile ? h->any_tile : h->any_rows) = true;
    const int ncr = nc[0] * nc[1] * nc[2];
    std::vector<std::pair<uint64_t, int>> key(ncr);
    for (int z = 0; z < nc[2]; ++z)
      for (int y = 0; y < nc[1]; ++y)
        for (int x = 0; x < nc[0]; ++x) {
          int lin = (z * nc[1] + y) * nc[0] + x;
          key[lin] = {morton3(x, y, z), lin};
        }
    std::sort(key.begin(), key.end());
    rank_all.resize(h->ncell + ncr);
    xyz_all.resize(h->ncell + ncr);
    crep_all.resize(h->ncell + ncr);
    for (int k = 0; k < ncr; ++k) {
      int lin = key[k].second, x = lin % nc[0], y = (lin / nc[0]) % nc[1], z = lin / (nc[0] * nc[1]);
      rank_all[h->ncell + lin] = h->ncell + k;
      xyz_all[h->ncell + k] = x | (y << 10) | (z << 20);
      crep_all[h->ncell + k] = r;
    }
    h->ncell += ncr;
  }
  for (int d = 0; d < 3; ++d) {
    h->nc[d] = h->hrep[0].nc[d];
    h->cell_len[d] = h->rep_box[d] / h->nc[d];
  }
  if ((rc = dalloc(h, &h->rdev, nrep))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->rdev, h->hrep.data(), sizeof(RepDev) * nrep, hipMemcpyHostToDevice, h->stream));
  if ((rc = dalloc(h, &h->cell_rank, h->ncell))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->cell_rank, rank_all.data(), sizeof(int) * h->ncell, hipMemcpyHostToDevice, h->stream));
  if ((rc = dalloc(h, &h->cell_xyz, h->ncell))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->cell_xyz, xyz_all.data(), sizeof(int) * h->ncell, hipMemcpyHostToDevice, h->stream));
  if (nrep > 1) {
    if ((rc = dalloc(h, &h->cell_rep, h->ncell))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->cell_rep, crep_all.data(), sizeof(int) * h->ncell, hipMemcpyHostToDevice, h->stream));
  }
  // rows of every group: replica r owns rows [row0[r], row0[r+1])
  std::vector<int> row0((size_t)(PFM_MAXG + 1) * (nrep + 1), 0);
  for (int r = 0; r < nrep; ++r) row0[r + 1] = h->rep_start[r] + h->rep_n[r];
  for (int g = 1; g <= PFM_MAXG; ++g) {
    h->rep_row0[g].assign(nrep + 1, 0);
    for (int r = 0; r < nrep; ++r) h->rep_row0[g][r + 1] = h->rep_row0[g][r] + gcnt[g][r];
    std::copy(h->rep_row0[g].begin(), h->rep_row0[g].end(), row0.begin() + (size_t)g * (nrep + 1));
  }
  if ((rc = dalloc(h, &h->d_rep_row0, row0.size()))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->d_rep_row0, row0.data(), sizeof(int) * row0.size(), hipMemcpyHostToDevice, h->stream));
  // counts, fill counters, then each cell's union of group bits
  if ((rc = dalloc(h, &h->cell_count, 4 * (size_t)h->ncell))) return rc;
  h->cell_fill = h->cell_count + h->ncell;
  h->cell_grp = (uint32_t*)(h->cell_count + 2 * (size_t)h->ncell);
  h->cell_tmp = h->cell_count + 3 * (size_t)h->ncell;
  if ((rc = dalloc(h, &h->cell_start, h->ncell))) return rc;
  // group rows
  for (int g = 1; g <= PFM_MAXG; ++g) {
    if (!h->grp.used[g]) continue;
    if ((rc = dalloc(h, &h->grp.rows[g], h->grp.size[g]))) return rc;
    if ((rc = dalloc(h, &h->grp.row_of[g], h->N))) return rc;
  }
  // list storage
  for (auto& it : h->inter) {
    for (int k = 0; k < it.nl_n; ++k) {
      ListDev& L = it.nl[k];
      {
        // atoms per cell over the replicas that hold rows of the list
        double atoms = 0.0, cells = 0.0;
        for (int r = 0; r < nrep; ++r) {
          if (gcnt[L.g1][r] == 0) continue;
          atoms += h->rep_n[r];
          cells += (double)h->hrep[r].nc[0] * h->hrep[r].nc[1] * h->hrep[r].nc[2];
        }
        L.sparse_cells = cells > 0.0 && atoms < 14.0 * cells;
      }
      // rows of the pair lists start on 16 bytes and end 32 entries early
      // enough that the passes' four-entry loads stay inside the array; the
      // graphene nearest-neighbour rows (k == 2) are three entries wide
      L.stride = k == 2 ? L.nmax : (L.nmax + 3) / 4 * 4;
      if ((rc = dalloc(h, &L.nnum, L.nrows))) return rc;
      if ((rc = dalloc(h, &L.nbr, (size_t)L.nrows * L.stride + 32))) return rc;
      if (it.kind == PFM_TB || it.kind == PFM_REBOSC || k == 2 || (it.kind == PFM_SUR && k == 1))
        if ((rc = dalloc(h, &L.rev, (size_t)L.nrows * L.stride))) return rc;
    }
    it.e_rows = it.nl[0].nrows;
    if ((rc = dalloc(h, &it.e_row, it.e_rows))) return rc;
    if (it.kind == PFM_LJ1G) {
      if (h->N > 0x3fffffff) return fail(h, PFM_E_INVALID, "error: lj1g lists hold atom indices below 2^30");
      if ((rc = dalloc(h, &it.lc, h->N))) return rc;
    }
    if (it.kind == PFM_SUR) {
      if ((rc = sur_alloc(h, it))) return rc;
      if (!it.sur_theta.empty()) {
        int ii = (int)(&it - h->inter.data()) + 1;
        if ((rc = pfm_sur_set_params(h, ii, it.sur_theta.data(), (int)it.sur_theta.size()))) return rc;
      }
    }
    if (it.kind == PFM_TB) {
      size_t ne = (size_t)it.nl[0].nrows * it.nl[0].stride;
      if ((rc = dalloc(h, &it.B, ne))) return rc;
      if ((rc = dalloc(h, &it.Bp, 2 * ne))) return rc;
    }
    if (is_cos(it.kind)) {
      if ((rc = dalloc(h, &it.gnorm, h->N))) return rc;
      if ((rc = dalloc(h, &it.gC, 3 * (size_t)h->N))) return rc;
      if ((rc = dalloc(h, &it.gnf, h->N))) return rc;
      if ((rc = dalloc(h, &it.mrows, (size_t)it.nl[1].nrows + 1))) return rc;
      std::vector<char> in1(h->N, 0);
      for (int x : h->grp.idx[it.nl[0].g1]) in1[x] = 1;
      it.rows_disjoint = true;
      for (int x : h->grp.idx[it.nl[0].g2])
        if (in1[x]) { it.rows_disjoint = false; break; }
    }
  }
  // pair searches (ListDev::pair_*): lj / lj1g lists in build order, each
  // paired with the first later unpaired one of the same rows, cutoff and
  // period when both are searched one wave per cell with over eight entries
  {
    struct Cand { int ii, k; };
    std::vector<Cand> order;
    for (size_t ii = 0; ii < h->inter.size(); ++ii) {
      Inter& it = h->inter[ii];
      if (it.kind != PFM_LJ && it.kind != PFM_LJ1G) continue;
      order.push_back({(int)ii, 0});
      if (has_converse(it.kind)) order.push_back({(int)ii, 1});
    }
    auto rc_of = [&](const Cand& c) { Inter& it = h->inter[c.ii]; return c.k == 1 ? it.nl[0].rcut : it.nl[c.k].rcut; };
    auto per_of = [&](const Cand& c) { Inter& it = h->inter[c.ii]; return c.k == 1 ? it.nl[0].period : it.nl[c.k].period; };
    for (auto& c : order) {
      ListDev& L = h->inter[c.ii].nl[c.k];
      L.pair_ii = L.pair_k = -1;
      L.pair_lead = L.pair_sparse = false;
      L.done_step = -1;
      L.quad_b1i = L.quad_b1k = L.quad_b2i = L.quad_b2k = -1;
      L.quad_done = -1;
    }
    for (size_t x = 0; x < order.size(); ++x) {
      ListDev& A = h->inter[order[x].ii].nl[order[x].k];
      if (A.pair_ii >= 0 || !A.sparse_cells || A.nmax <= 8) continue;
      for (size_t y = x + 1; y < order.size(); ++y) {
        ListDev& B = h->inter[order[y].ii].nl[order[y].k];
        if (B.pair_ii >= 0 || !B.sparse_cells || B.nmax <= 8 || B.g1 != A.g1 || rc_of(order[y]) != rc_of(order[x]) ||
            per_of(order[y]) != per_of(order[x]))
          continue;
        A.pair_ii = order[y].ii; A.pair_k = order[y].k; A.pair_lead = true;
        B.pair_ii = order[x].ii; B.pair_k = order[x].k;
        std::vector<char> in(h->N, 0);
        long nu = 0;
        for (int v : h->grp.idx[A.g2]) if (!in[v]) { in[v] = 1; ++nu; }
        for (int v : h->grp.idx[B.g2]) if (!in[v]) { in[v] = 1; ++nu; }
        A.pair_sparse = 2 * nu < (long)h->N;
        break;
      }
    }
    // quads: two pairs (leads P and Q, P first) whose row groups share no
    // atom and whose lists take the same two candidate groups, with the same
    // cutoff, period and staging
    std::vector<Cand> leads;
    for (auto& c : order)
      if (h->inter[c.ii].nl[c.k].pair_lead) leads.push_back(c);
    for (size_t x = 0; x < leads.size(); ++x) {
      ListDev& P = h->inter[leads[x].ii].nl[leads[x].k];
      if (!P.pair_lead || P.quad_b1i >= 0) continue;
      ListDev& Pp = h->inter[P.pair_ii].nl[P.pair_k];
      for (size_t y = x + 1; y < leads.size(); ++y) {
        ListDev& Q = h->inter[leads[y].ii].nl[leads[y].k];
        if (!Q.pair_lead || Q.quad_b1i >= 0) continue;
        ListDev& Qp = h->inter[Q.pair_ii].nl[Q.pair_k];
        if (rc_of(leads[y]) != rc_of(leads[x]) || per_of(leads[y]) != per_of(leads[x]) ||
            Q.pair_sparse != P.pair_sparse || P.g2 == Pp.g2)
          continue;
        // Q's lists: the one on P's candidate group and the one on P's partner's
        Cand q1{leads[y].ii, leads[y].k}, q2{Q.pair_ii, Q.pair_k};
        if (Q.g2 == Pp.g2 && Qp.g2 == P.g2) std::swap(q1, q2);
        else if (!(Q.g2 == P.g2 && Qp.g2 == Pp.g2)) continue;
        std::vector<char> in(h->N, 0);
        bool disjoint = true;
        for (int v : h->grp.idx[P.g1]) in[v] = 1;
        for (int v : h->grp.idx[Q.g1]) disjoint = disjoint && !in[v];
        if (!disjoint) continue;
        P.quad_b1i = q1.ii; P.quad_b1k = q1.k;
        P.quad_b2i = q2.ii; P.quad_b2k = q2.k;
        Q.pair_lead = false;  // built in P's pass
        break;
      }
    }
  }
  // graphene on a metal: exactly rjl, tb and a ljc/morsec fed by that tb
  // list, with the rjl rows, the graphene rows and the metal rows of
  // ljc/morsec sharing no atom (the passes of each side may then share
  // launches, every force being summed by one row)
  h->gm = pfm_ctx::GmPlan{};
  if (h->inter.size() == 3) {
    int ir = -1, itb = -1, ic = -1;
    for (int k = 0; k < 3; ++k) {
      const int kind = h->inter[k].kind;
      if (kind == PFM_RJL) ir = k;
      else if (kind == PFM_TB) itb = k;
      else if (is_cos(kind)) ic = k;
    }
    if (ir >= 0 && itb >= 0 && ic >= 0 && h->inter[ic].nn_from_tb == itb && h->inter[ic].rows_disjoint) {
      std::vector<char> inC(h->N, 0);
      for (int x : h->grp.idx[h->inter[itb].nl[0].g1]) inC[x] = 1;
      bool ok = true;
      for (int x : h->grp.idx[h->inter[ir].nl[0].g1]) ok = ok && !inC[x];
      for (int x : h->grp.idx[h->inter[ic].nl[1].g1]) ok = ok && !inC[x];
      if (ok) {
        h->gm = pfm_ctx::GmPlan{true, ir, itb, ic, true};
        std::vector<char> inG(h->N, 0);
        for (int x : h->grp.idx[h->inter[ic].nl[0].g1]) inG[x] = 1;
        for (int x : h->grp.idx[h->inter[ir].nl[0].g1]) h->gm.first_ok = h->gm.first_ok && !inG[x];
      }
    }
  }
  // group-local index by original atom, for lj1g and rebosc
  for (auto& it : h->inter) {
    if (it.kind != PFM_LJ1G && it.kind != PFM_REBOSC) continue;
    int g = it.nl[0].g2;
    if (h->grp.loc[g]) continue;
    std::vector<int> loc(h->N, -1);
    for (size_t k = 0; k < h->grp.idx[g].size(); ++k) loc[h->grp.idx[g][k]] = (int)k;
    if ((rc = dalloc(h, &h->grp.loc[g], h->N))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->grp.loc[g], loc.data(), sizeof(int) * h->N, hipMemcpyHostToDevice, h->stream));
  }
  // group bit masks: every member active until the first step applies the changes
  std::vector<uint32_t> mask(h->N, 0);
  for (int g = 1; g <= PFM_MAXG; ++g)
    for (int a : h->grp.idx[g]) {
      mask[a] |= 1u << (g - 1);
      if (h->grp.shadow[g]) mask[a] |= 1u << (h->grp.shadow[g] - 1);
    }
  for (int g = 1; g <= PFM_MAXG; ++g) {
    if (!h->grp.shadow[g]) continue;
    std::vector<int> rank(h->N, -1);
    if (nrep > 1) {
      // position among the members of the atom's own replica
      std::vector<int> seen(nrep, 0);
      for (size_t k = 0; k < h->grp.idx[g].size(); ++k) {
        const int a = h->grp.idx[g][k];
        rank[a] = seen[rep_of[a]]++;
      }
      if ((rc = dalloc(h, &h->grp.d_rn[g], nrep))) return rc;
      HIPCHK(h, hipMemcpyAsync(h->grp.d_rn[g], h->grp.rn[g].data(), sizeof(int) * nrep, hipMemcpyHostToDevice,
                               h->stream));
    } else {
      for (size_t k = 0; k < h->grp.idx[g].size(); ++k) rank[h->grp.idx[g][k]] = (int)k;
    }
    if ((rc = dalloc(h, &h->grp.rank[g], h->N))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->grp.rank[g], rank.data(), sizeof(int) * h->N, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipMemcpyAsync(h->mask, mask.data(), sizeof(uint32_t) * h->N, hipMemcpyHostToDevice, h->stream));
  // chains
  if (!h->nhc.empty()) {
    std::vector<NhcDev> cs(h->nhc.size());
    for (size_t c = 0; c < h->nhc.size(); ++c) {
      NhcDev& d = cs[c];
      memset(&d, 0, sizeof d);
      d.M = h->nhc[c].M;
      d.L = (double)h->nhc[c].L;
      d.T = h->nhc[c].T;
      d.kt = PFM_KB * h->nhc[c].T;
      d.q[0] = h->nhc[c].q1;
      for (int i = 1; i < d.M; ++i) d.q[i] = d.q[0] / (3.0 * h->nhc[c].L);
      d.s = 1.0;
      d.bit = 1u << (h->nhc[c].group - 1);
    }
    if ((rc = dalloc(h, (NhcDev**)&h->nhc_state, cs.size()))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->nhc_state, cs.data(), sizeof(NhcDev) * cs.size(), hipMemcpyHostToDevice, h->stream));
    h->e_nhc.assign(h->nhc.size(), 0.0);
  }
  return PFM_OK;
}

static int check_err(pfm_ctx* h, int s1) {
  HIPCHK(h, hipMemcpyAsync(h->herr, h->derr, sizeof(DevErr), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  DevErr e = *h->herr;
  if (e.code == 0) return PFM_OK;
  if (e.code == PFM_E_OUT_OF_CELL && e.step > s1) {
    // found by the last drift: the reference stops at the next step's check
    HIPCHK(h, hipMemsetAsync(h->derr, 0, sizeof(DevErr), h->stream));
    return PFM_OK;
  }
  HIPCHK(h, hipMemsetAsync(h->derr, 0, sizeof(DevErr), h->stream));
  char buf[256];
  switch (e.code) {
    case PFM_E_OVERFLOW: snprintf(buf, sizeof buf, "error: too many neighbours %d %d (step %d)", e.atom, e.count, e.step); break;
    case PFM_E_OUT_OF_CELL:
      snprintf(buf, sizeof buf, "%d particle out of cell %.10g %.10g %.10g (step %d)", e.atom, e.x, e.y, e.z, e.step);
      break;
    case PFM_E_GRAPHENE_NN:
      if (e.count < 0) snprintf(buf, sizeof buf, "l1>3 (atom %d, step %d)", e.atom, e.step);
      else if (e.count > 3) snprintf(buf, sizeof buf, "error: too many gr nearest neibs %d %d (step %d)", e.atom, e.count, e.step);
      else snprintf(buf, sizeof buf, "error: not enough gr nearest neibs %d %d (step %d)", e.atom, e.count, e.step);
      break;
    default: snprintf(buf, sizeof buf, "error: device error %d", e.code);
  }
  return fail(h, e.code, buf);
}

static IntP int_params(pfm_ctx* h, double dt) {
  IntP P;
  memset(&P, 0, sizeof P);
  P.ts0 = dt;
  P.ts1 = dt / 2.0;
  P.bxyz = h->roles[1] ? 1u << (h->roles[1] - 1) : 0;
  P.bz = h->roles[2] ? 1u << (h->roles[2] - 1) : 0;
  for (int d = 0; d < 3; ++d) P.box[d] = h->box[d];
  P.nchain = (int)h->nhc.size();
  for (int c = 0; c < P.nchain; ++c) P.chain_bit[c] = 1u << (h->nhc[c].group - 1);
  P.chain_s = h->red_out + 16;
  P.zlo = 0.8 * h->box[2];
  P.zhi = 0.9 * h->box[2];
  return P;
}

// The top of the step at which an nvms run ends: md() runs check_positions
// and invert_z_velocities before its exit test (md_simulation.f90:139-146),
// so the state it leaves has the wall applied and an atom the last drift put
// out of the cell still stops the run.
static int exit_step_top(pfm_ctx* h, int s) {
  const double zlo = 0.8 * h->box[2], zhi = 0.9 * h->box[2];
  PFM_LAUNCH(h, "check", k_check, (h->N + 255) / 256, 256, h->N, h->pos4, h->box[0], h->box[1], h->box[2], h->derr,
             h->orig, s, h->invert_z, h->vel4, zlo, zhi, h->rep, h->rdev);
  return check_err(h, s);
}

static MaskQ chain_masks(pfm_ctx* h) {
  MaskQ Q;
  memset(&Q, 0, sizeof Q);
  Q.nq = (int)h->nhc.size();
  for (int c = 0; c < Q.nq; ++c) Q.bit[c] = 1u << (h->nhc[c].group - 1);
  return Q;
}

// first NHC half step of a step: kinetic energies, then the chains
static void nhc_first_half(pfm_ctx* h, double dt) {
  int nb = h->red_blocks;
  MaskQ Q = chain_masks(h);
  ChainArgs A{(NhcDev*)h->nhc_state, h->red_out + 16, dt / 2.0, dt / 4.0, dt / 8.0, 0};
  PFM_LAUNCH(h, "nhc", k_ke_part, nb, RED_THREADS, h->N, h->vel4, h->mask, Q, h->red_part);
  PFM_LAUNCH(h, "nhc", k_nhc_chain, 1, RED_THREADS, Q.nq, A, h->red_part, nb);
}

// per-block partials of the interaction energies and the all_moving kinetic
// energy into red_part, and of the out block into `opart` unless it is null
static void energy_parts(pfm_ctx* h, double* opart) {
  RecPartArgs A;
  memset(&A, 0, sizeof A);
  A.ni = (int)h->inter.size();
  A.N = h->N;
  for (int k = 0; k < A.ni; ++k) {
    A.e_row[k] = h->inter[k].e_row;
    A.e_rows[k] = h->inter[k].e_rows;
  }
  A.pos = h->pos4;
  A.vel = h->vel4;
  A.frc = h->frc4;
  A.mask = h->mask;
  A.ke_bit = 1u << (h->roles[0] - 1);
  A.all_bit = 1u << (h->roles[3] - 1);
  A.epart = h->red_part;
  A.opart = opart;
  PFM_LAUNCH(h, "energy", k_rec_part, h->red_blocks, RED_THREADS, A);
}

// Energies of the step: per-interaction sums of the per-row energies and the
// all_moving group's kinetic energy; one blocking read-back.
// arguments of k_rep_energy over the handle's replicas (out left to the caller)
static RepEnergyArgs rep_energy_args(pfm_ctx* h) {
  RepEnergyArgs A;
  memset(&A, 0, sizeof A);
  const int ni = (int)h->inter.size();
  A.ni = ni;
  A.nrep = h->nrep;
  for (int k = 0; k < ni; ++k) {
    A.e_row[k] = h->inter[k].e_row;
    A.g1[k] = h->inter[k].nl[0].g1;
  }
  A.row0 = h->d_rep_row0;
  A.vel = h->vel4;
  A.mask = h->mask;
  A.bit = 1u << (h->roles[0] - 1);
  return A;
}

// per-replica energies from the [nrep][ni + 1] rows of k_rep_energy
static void set_rep_energies(pfm_ctx* h, const double* o) {
  const int nrep = h->nrep, ni = (int)h->inter.size();
  h->rep_e.assign(o, o + (size_t)nrep * (ni + 1));
  h->rep_pe.assign(nrep, 0.0);
  h->rep_ke.assign(nrep, 0.0);
  h->rep_temp.assign(nrep, 0.0);
  const std::vector<int>& m0 = h->rep_row0[h->roles[0]];
  const std::vector<int>& act = h->grp.rn[h->roles[0]];  // active counts of a growing batch
  for (int r = 0; r < nrep; ++r) {
    double p = 0.0;
    for (int k = 0; k < ni; ++k) p += o[(size_t)r * (ni + 1) + k];
    h->rep_pe[r] = p;
    h->rep_ke[r] = o[(size_t)r * (ni + 1) + ni];
    int nm = act.empty() ? m0[r + 1] - m0[r] : act[r];
    h->rep_temp[r] = nm > 0 ? 2 * h->rep_ke[r] / PFM_KB / (3 * nm) : 0.0;
  }
}

static int energies(pfm_ctx* h, bool nvt) {
  int nb = h->red_blocks;
  int ni = (int)h->inter.size();
  energy_parts(h, nullptr);
  PFM_LAUNCH(h, "energy", k_reduce_rows, 1, RED_THREADS, ni + 1, nb, h->red_part, h->red_out + 32);
  HIPCHK(h, hipMemcpyAsync(h->h_red, h->red_out, sizeof(double) * (32 + PFM_MAXI + 1), hipMemcpyDeviceToHost,
                           h->stream));
  std::vector<NhcDev> cs(nvt ? h->nhc.size() : 0);
  if (!cs.empty())
    HIPCHK(h, hipMemcpyAsync(cs.data(), h->nhc_state, sizeof(NhcDev) * cs.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double pe = 0.0;
  for (int k = 0; k < ni; ++k) {
    h->inter[k].energy = h->h_red[32 + k];
    pe += h->inter[k].energy;
  }
  h->pe = pe;
  h->ke = h->h_red[32 + ni];
  h->temp = 2 * h->ke / PFM_KB / (3 * h->grp.n[h->roles[0]]);
  if (h->nrep > 1) {
    const int nrep = h->nrep;
    if (!h->d_rep_out) {
      int rc;
      if ((rc = dalloc(h, &h->d_rep_out, (size_t)nrep * (ni + 1)))) return rc;
    }
    RepEnergyArgs A = rep_energy_args(h);
    A.out = h->d_rep_out;
    PFM_LAUNCH(h, "energy", k_rep_energy, nrep, RED_THREADS, A);
    std::vector<double> o((size_t)nrep * (ni + 1));
    HIPCHK(h, hipMemcpyAsync(o.data(), h->d_rep_out, sizeof(double) * o.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    set_rep_energies(h, o.data());
  }
  for (size_t c = 0; c < cs.size(); ++c) {
    // calculate_nose_hoover_chain_energy (md_integrators.f90:247-260)
    const NhcDev& d = cs[c];
    double e = d.q[0] / 2 * d.v[0] * d.v[0] + 3.0 * d.L * d.kt * d.x[0];
    for (int i = 1; i < d.M; ++i) e = e + d.q[i] / 2 * d.v[i] * d.v[i] + d.kt * d.x[i];
    h->e_nhc[c] = e;
  }
  return PFM_OK;
}

// ------------------------------------------------------------ nvms blocks

// the state of one step into ring slot k, or back from it
__global__ void k_spec_copy(int N, int back, double4* __restrict__ pos, double4* __restrict__ vel,
                            double4* __restrict__ frc, uint32_t* __restrict__ mask, int* __restrict__ orig,
                            double4* __restrict__ rpos, double4* __restrict__ rvel, double4* __restrict__ rfrc,
                            uint32_t* __restrict__ rmask, int* __restrict__ rorig) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (back) {
    pos[i] = rpos[i]; vel[i] = rvel[i]; frc[i] = rfrc[i]; mask[i] = rmask[i]; orig[i] = rorig[i];
  } else {
    rpos[i] = pos[i]; rvel[i] = vel[i]; rfrc[i] = frc[i]; rmask[i] = mask[i]; rorig[i] = orig[i];
  }
}

static int spec_alloc(pfm_ctx* h, int k) {
  if (h->spec_k >= k) return PFM_OK;
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(h->xfer);
  F(h->spec_e); F(h->sp_pos); F(h->sp_vel); F(h->sp_frc); F(h->sp_mask); F(h->sp_orig);
  h->spec_e = nullptr; h->sp_pos = h->sp_vel = h->sp_frc = nullptr; h->sp_mask = nullptr; h->sp_orig = nullptr;
  h->spec_k = 0;
  const size_t n = (size_t)k * h->N;
  const size_t nr = h->nrep > 1 ? (size_t)h->nrep : 0;  // batches: per-replica rows after the whole-system rows
  int rc;
  if ((rc = dalloc(h, &h->spec_e, (size_t)k * (h->inter.size() + 1) * (1 + nr))) || (rc = dalloc(h, &h->sp_pos, n)) ||
      (rc = dalloc(h, &h->sp_vel, n)) || (rc = dalloc(h, &h->sp_frc, n)) || (rc = dalloc(h, &h->sp_mask, n)) ||
      (rc = dalloc(h, &h->sp_orig, n)))
    return rc;
  h->spec_k = k;
  return PFM_OK;
}

// the energies of the step into slot k (the sums of energies(); for a batch
// also each replica's, after the spec_k whole-system rows) and its state
static void spec_record(pfm_ctx* h, int k) {
  const int nb = h->red_blocks, ni = (int)h->inter.size(), N = h->N;
  energy_parts(h, nullptr);
  PFM_LAUNCH(h, "energy", k_reduce_rows, 1, RED_THREADS, ni + 1, nb, h->red_part, h->spec_e + (size_t)k * (ni + 1));
  if (h->nrep > 1) {
    RepEnergyArgs A = rep_energy_args(h);
    A.out = h->spec_e + (size_t)h->spec_k * (ni + 1) + (size_t)k * h->nrep * (ni + 1);
    PFM_LAUNCH(h, "energy", k_rep_energy, h->nrep, RED_THREADS, A);
  }
  const size_t o = (size_t)k * N;
  PFM_LAUNCH(h, "spec", k_spec_copy, (N + 255) / 256, 256, N, 0, h->pos4, h->vel4, h->frc4, h->mask, h->orig,
             h->sp_pos + o, h->sp_vel + o, h->sp_frc + o, h->sp_mask + o, h->sp_orig + o);
}

// rebuild every list due at step s, in settings order; re-sort the atoms
// first when all of them are due (step 0 always)
// list k of interaction ii, with its pair when it leads one; a list its
// lead built at this step is skipped
static int build_or_pair(pfm_ctx* h, int ii, int k, int s) {
  ListDev& L = h->inter[ii].nl[k];
  if (L.pair_ii >= 0 && !L.pair_lead && (L.done_step == s || L.quad_done == s)) return PFM_OK;
  if (L.pair_lead) {
    h->inter[L.pair_ii].nl[L.pair_k].done_step = s;
    if (L.quad_b1i >= 0) {
      h->inter[L.quad_b1i].nl[L.quad_b1k].quad_done = s;
      h->inter[L.quad_b2i].nl[L.quad_b2k].quad_done = s;
      return launch_build_quad(h, ii, k, L.pair_ii, L.pair_k, L.quad_b1i, L.quad_b1k, L.quad_b2i, L.quad_b2k, s);
    }
    return launch_build_pair(h, ii, k, L.pair_ii, L.pair_k, s);
  }
  return launch_build_list(h, ii, k, s);
}

static int update_lists(pfm_ctx* h, int s) {
  bool all_due = true;
  for (auto& it : h->inter)
    for (int k = 0; k < it.nl_n; ++k) {
      int per = (k == 1 || (k == 2 && it.nn_from_tb >= 0)) ? -1 : it.nl[k].period;
      if (k == 2 && it.nn_from_tb >= 0) per = h->inter[it.nn_from_tb].nl[0].period;
      if (k == 1 && has_converse(it.kind)) per = it.nl[0].period;
      if (per > 0 && s % per != 0) all_due = false;
    }
  int rc;
  if (all_due || !h->sorted_once || h->need_sort) {
    h->need_sort = false;
    if ((rc = launch_sort_and_bin(h, s, true))) return rc;
    h->sorted_once = true;
    for (auto& it : h->inter)
      for (int k = 0; k < it.nl_n; ++k) it.nl[k].built = false;
  } else {
    // some lists are due, others keep their entries: bin the atoms at their
    // current positions without moving them
    bool any_due = false;
    for (auto& it : h->inter) {
      any_due = any_due || s % it.nl[0].period == 0;
      if (is_cos(it.kind) && it.nn_from_tb < 0) any_due = any_due || s % it.nl[2].period == 0;
    }
    if (any_due && (rc = launch_sort_and_bin(h, s, false))) return rc;
  }
  for (size_t ii = 0; ii < h->inter.size(); ++ii) {
    Inter& it = h->inter[ii];
    bool due = s % it.nl[0].period == 0 || !it.nl[0].built;
    if (due) {
      if ((rc = build_or_pair(h, (int)ii, 0, s))) return rc;
      if (it.kind == PFM_SUR) launch_sur_converse(h, it, s);
      else if (has_converse(it.kind))
        if ((rc = build_or_pair(h, (int)ii, 1, s))) return rc;
    }
    if (is_cos(it.kind)) {
      if (it.nn_from_tb >= 0) {
        Inter& tb = h->inter[it.nn_from_tb];
        if (s % tb.nl[0].period == 0 || !it.nl[2].built)
          if ((rc = launch_nn_from_tb(h, (int)ii, s))) return rc;
      } else if (s % it.nl[2].period == 0 || !it.nl[2].built) {
        if ((rc = launch_build_list(h, (int)ii, 2, s))) return rc;
      }
    }
  }
  return PFM_OK;
}

static int zero_momentum(pfm_ctx* h) {
  int nb = h->red_blocks, g = h->roles[3];
  if (g == 0 || h->grp.n[g] == 0) return PFM_OK;
  uint32_t bit = 1u << (g - 1);
  if (h->nrep > 1) {
    PFM_LAUNCH(h, "zero_momentum", k_rep_zero_mom, h->nrep, RED_THREADS, h->nrep, h->d_rep_row0, h->vel4, h->mask, bit);
    return PFM_OK;
  }
  PFM_LAUNCH(h, "zero_momentum", k_mom_part, nb, RED_THREADS, h->N, h->vel4, h->mask, bit, h->red_part);
  PFM_LAUNCH(h, "zero_momentum", k_reduce_rows, 1, RED_THREADS, 4, nb, h->red_part, h->red_out);
  PFM_LAUNCH(h, "zero_momentum", k_sub_mcv, (h->N + 255) / 256, 256, h->N, h->vel4, h->mask, bit, h->red_out);
  return PFM_OK;
}

// ------------------------------------------------------- captured blocks

__global__ void k_set_base(DevErr* e, int base) {
  if (threadIdx.x == 0) e->base = base;
}

static long gcd_l(long a, long b) { return b ? gcd_l(b, a % b) : a; }

// captured blocks kept per handle: both re-sort buffer sets times the block
// variants a schedule of zero-momentum and output steps produces
static const size_t GRAPH_CACHE = 16;

// steps that take output records (pfm_set_records)
static bool rec_on(const pfm_ctx* h) { return h->rec.fn && (!h->rec.eper.empty() || !h->rec.xper.empty()); }
static bool due(const std::vector<int>& per, int s) {
  for (int p : per)
    if (s % p == 0) return true;
  return false;
}
static bool rec_e_at(const pfm_ctx* h, int s) { return rec_on(h) && due(h->rec.eper, s); }
static bool rec_x_at(const pfm_ctx* h, int s) { return rec_on(h) && due(h->rec.xper, s); }

// length of the repeating schedule of list rebuilds and zero_momentum, or 0
// when it is longer than 64 steps
static int graph_block(pfm_ctx* h) {
  long p = 1;
  for (auto& it : h->inter)
    for (int k = 0; k < it.nl_n; ++k) {
      long q = it.nl[k].period;
      p = p / gcd_l(p, q) * q;
      if (p > 64) return 0;
    }
  if (h->zm_period <= 64) p = p / gcd_l(p, h->zm_period) * h->zm_period;
  return p > 64 ? 0 : (int)p;
}

static std::vector<uint64_t> graph_buffers(pfm_ctx* h) {
  return {(uint64_t)h->pos4, (uint64_t)h->vel4, (uint64_t)h->frc4, (uint64_t)h->mask, (uint64_t)h->orig,
          (uint64_t)h->rep, (uint64_t)h->pos4b, (uint64_t)h->vel4b, (uint64_t)h->frc4b, (uint64_t)h->maskb,
          (uint64_t)h->origb, (uint64_t)h->repb, (uint64_t)h->bin};
}

static void set_graph_buffers(pfm_ctx* h, const std::vector<uint64_t>& b) {
  h->pos4 = (double4*)b[0]; h->vel4 = (double4*)b[1]; h->frc4 = (double4*)b[2];
  h->mask = (uint32_t*)b[3]; h->orig = (int*)b[4]; h->rep = (int*)b[5];
  h->pos4b = (double4*)b[6]; h->vel4b = (double4*)b[7]; h->frc4b = (double4*)b[8];
  h->maskb = (uint32_t*)b[9]; h->origb = (int*)b[10]; h->repb = (int*)b[11];
  h->bin = (const int*)b[12];
}

// what the launches of a block starting at step s depend on
static std::vector<uint64_t> graph_key(pfm_ctx* h, int integ, double dt, int blk, bool first_writes, int s) {
  uint64_t dtb;
  memcpy(&dtb, &dt, 8);
  std::vector<uint64_t> k = graph_buffers(h);
  bool zm = false;
  for (int t = s; t < s + blk; ++t) zm = zm || t % h->zm_period == 0;
  uint64_t em = 0, xm = 0;
  for (int t = s; t < s + blk; ++t) {
    em |= (uint64_t)rec_e_at(h, t) << (t - s);
    xm |= (uint64_t)rec_x_at(h, t) << (t - s);
  }
  k.insert(k.end(), {(uint64_t)integ, dtb, (uint64_t)blk, (uint64_t)first_writes, (uint64_t)zm,
                     (uint64_t)h->invert_z, (uint64_t)h->lanes, (uint64_t)h->block, (uint64_t)h->fuse, em, xm});
  for (const auto& it : h->inter) k.push_back((uint64_t)it.lanes);
  return k;
}

// Graphs are launched asynchronously: one may still be running on the
// handle's stream, so the stream is drained before any executable graph is
// destroyed (the runtime and any tool hooked into it may still read the
// graph's launch records until its last kernel has finished).
static void clear_graphs(pfm_ctx* h) {
  if (h->graphs.empty()) return;
  (void)hipStreamSynchronize(h->stream);
  for (auto& e : h->graphs) (void)hipGraphExecDestroy(e.exec);
  h->graphs.clear();
}

// ------------------------------------------------------------- records

static_assert(sizeof(NhcDev) % sizeof(double) == 0, "chain state copied as doubles");
static const int REC_ECAP = 256;  // energy slots
#define REC_OUT 11                // out-block values per energy record


static void rec_free(pfm_ctx* h) {
  auto& R = h->rec;
  // record kernels may still write the pinned slots
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (R.d_x) (void)hipFree(R.d_x);
  if (R.d_part) (void)hipFree(R.d_part);
  if (R.d_ctr) (void)hipFree(R.d_ctr);
  if (R.h_e) (void)hipHostFree(R.h_e);
  if (R.h_x) (void)hipHostFree(R.h_x);
  R.d_x = R.d_part = R.h_e = R.h_x = nullptr;
  R.d_ctr = nullptr;
  R.ecap = R.xcap = R.ewidth = 0;
  for (auto& p : R.q) R.free_ev.push_back(p.done);
  R.q.clear();
  R.pending_e = R.pending_x = 0;
}

// device and pinned host rings for the current interactions and chains
static int rec_prepare(pfm_ctx* h) {
  auto& R = h->rec;
  const int ni = (int)h->inter.size(), nq = (int)h->nhc.size();
  const int width = ni + 1 + REC_OUT + nq * (int)(sizeof(NhcDev) / sizeof(double));
  const int want_x = R.xper.empty() ? 0 : R.xslots;
  if (R.d_ctr && R.ewidth == width && R.xcap == want_x) return PFM_OK;
  rec_free(h);
  int rc;
  R.ewidth = width;
  R.ecap = REC_ECAP;
  R.xcap = want_x;
  if ((rc = dalloc(h, &R.d_part, (size_t)(2 * (REC_OUT - 1) + 1) * h->red_blocks)) || (rc = dalloc(h, &R.d_ctr, 4)))
    return rc;
  // the record kernels write the slots in pinned host memory themselves
  HIPCHK(h, hipHostMalloc((void**)&R.h_e, sizeof(double) * R.ecap * width, hipHostMallocCoherent));
  if (R.xcap > 0) {
    if ((rc = dalloc(h, &R.d_x, (size_t)6 * h->N))) return rc;
    HIPCHK(h, hipHostMalloc((void**)&R.h_x, sizeof(double) * (size_t)R.xcap * 6 * h->N, hipHostMallocCoherent));
  }
  return PFM_OK;
}

// launches that take the records of the current step (captured with it)
static void rec_launch(pfm_ctx* h, bool e, bool x) {
  auto& R = h->rec;
  const int nb = h->red_blocks, ni = (int)h->inter.size();
  if (e) energy_parts(h, R.d_part);
  if (x) {
    const size_t n2 = (size_t)3 * h->N;
    PFM_LAUNCH(h, "record", k_rec_frame, (h->N + 255) / 256, 256, h->N, h->pos4, h->vel4, h->orig, R.d_x);
    PFM_LAUNCH(h, "record", k_rec_frame_out, (int)std::min<size_t>(1024, (n2 + 255) / 256), 256, n2,
               (const double2*)R.d_x, R.h_x, R.xcap, R.d_ctr);
  }
  if (e)
    PFM_LAUNCH(h, "record", k_rec_energy, 1, RED_THREADS, ni + 1, nb, h->red_part, R.d_part, h->nhc_state,
               (int)(h->nhc.size() * sizeof(NhcDev) / sizeof(double)), R.h_e, R.ewidth, R.ecap, R.d_ctr, x ? 1 : 0);
  else
    PFM_LAUNCH(h, "record", k_rec_bump, 1, 64, R.d_ctr, 0, 1);
}

static double nhc_energy(const NhcDev& d) {
  // calculate_nose_hoover_chain_energy (md_integrators.f90:247-260)
  double e = d.q[0] / 2 * d.v[0] * d.v[0] + 3.0 * d.L * d.kt * d.x[0];
  for (int i = 1; i < d.M; ++i) e = e + d.q[i] / 2 * d.v[i] * d.v[i] + d.kt * d.x[i];
  return e;
}

// hand the oldest pending record to the callback; wait for its slot when
// `wait`, else only if it is written.  Returns 1 when nothing was handed on.
static int rec_deliver_one(pfm_ctx* h, bool wait, int* rc) {
  auto& R = h->rec;
  *rc = PFM_OK;
  if (R.q.empty()) return 1;
  pfm_ctx::Records::Pending P = R.q.front();
  if (wait) {
    hipError_t e = hipEventSynchronize(P.done);
    if (e != hipSuccess) {
      *rc = fail(h, PFM_E_HIP, std::string("error: HIP ") + hipGetErrorString(e) + " waiting for a record");
      return 1;
    }
  } else {
    hipError_t e = hipEventQuery(P.done);
    if (e == hipErrorNotReady) return 1;
    if (e != hipSuccess) {
      *rc = fail(h, PFM_E_HIP, std::string("error: HIP ") + hipGetErrorString(e) + " polling a record");
      return 1;
    }
  }
  R.q.pop_front();
  R.free_ev.push_back(P.done);
  pfm_record r;
  memset(&r, 0, sizeof r);
  r.step = P.step;
  r.integrator = P.integ;
  r.sim_time = P.time;
  r.group_n = P.group_n.data();
  const int ni = (int)h->inter.size(), nq = (int)h->nhc.size();
  std::vector<double> enhc(h->e_nhc);
  enhc.resize(nq, 0.0);
  if (P.eslot >= 0) {
    --R.pending_e;
    const double* e = R.h_e + (size_t)P.eslot * R.ewidth;
    r.has_energy = 1;
    double pe = 0.0;
    for (int k = 0; k < ni; ++k) pe += e[k];
    r.pe = pe;
    r.ke = e[ni];
    r.temp = 2 * r.ke / PFM_KB / (3 * P.group_n[h->roles[0]]);
    const double* o = e + ni + 1;
    const double tm = o[9];
    for (int d = 0; d < 3; ++d) {
      r.force_sum[d] = o[d];
      r.mass_center[d] = o[3 + d] / tm;
      r.mass_center_vel[d] = o[6 + d] / tm;
    }
    r.max_vel = std::sqrt(o[10]);
    if (P.integ == PFM_NVT) {
      const NhcDev* cs = (const NhcDev*)(e + ni + 1 + REC_OUT);
      for (int c = 0; c < nq; ++c) {
        NhcDev d;
        memcpy(&d, cs + c, sizeof d);
        enhc[c] = nhc_energy(d);
      }
    }
    r.e_inter = e;
    r.e_nhc = enhc.data();
    if (P.integ != PFM_NVMS) {
      // the record is the handle's latest energy evaluation (pfm_energies)
      for (int k = 0; k < ni; ++k) h->inter[k].energy = e[k];
      h->pe = pe;
      h->ke = r.ke;
      h->temp = r.temp;
      h->e_nhc = enhc;
    }
  }
  if (P.xslot >= 0) {
    --R.pending_x;
    r.has_frame = 1;
    r.pos = R.h_x + (size_t)P.xslot * 6 * h->N;
    r.vel = r.pos + 3 * (size_t)h->N;
  }
  int u = R.fn(R.user, &r);
  if (u != 0) {
    char buf[96];
    snprintf(buf, sizeof buf, "error: record callback returned %d at step %d", u, P.step);
    *rc = fail(h, PFM_E_INVALID, buf);
    return 1;
  }
  return 0;
}

// free slots for e energy records and x frames: hand on the oldest records
// (waiting for their slots) until they fit
static int rec_reserve(pfm_ctx* h, int e, int x) {
  auto& R = h->rec;
  int rc;
  while (R.pending_e + e > R.ecap || R.pending_x + x > R.xcap) {
    if (rec_deliver_one(h, true, &rc)) return rc ? rc : fail(h, PFM_E_INVALID, "error: record rings too small");
  }
  return PFM_OK;
}

// note the records whose launches were just queued on the handle's stream:
// an event after them marks when their pinned slots are written
struct RecNote {
  int step;
  double time;
  bool e, x;
};
static int rec_issue(pfm_ctx* h, int integ, const std::vector<RecNote>& notes) {
  auto& R = h->rec;
  if (notes.empty()) return PFM_OK;
  for (const RecNote& n : notes) {
    pfm_ctx::Records::Pending P{n.step, integ, n.time, -1, -1, nullptr, {}};
    P.group_n.assign(h->grp.n, h->grp.n + PFM_MAXG + 1);
    if (n.e) {
      P.eslot = (int)(R.ne++ % R.ecap);
      ++R.pending_e;
    }
    if (n.x) {
      P.xslot = (int)(R.nx++ % R.xcap);
      ++R.pending_x;
    }
    if (R.free_ev.empty()) {
      hipEvent_t ev;
      HIPCHK(h, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      R.free_ev.push_back(ev);
    }
    P.done = R.free_ev.back();
    R.free_ev.pop_back();
    HIPCHK(h, hipEventRecord(P.done, h->stream));
    R.q.push_back(P);
  }
  return PFM_OK;
}

// hand on every record that has arrived (all of them when `wait`)
static int rec_poll(pfm_ctx* h, bool wait) {
  int rc = PFM_OK;
  while (!rec_deliver_one(h, wait, &rc)) {
  }
  return rc;
}

// ---------------------------------------------------------------- C-ABI

extern "C" {

int pfm_create(pfm_ctx** out, int n, const double box[3], const double* pos, const double* vel, const double* mass,
               int precision) {
  return pfm_create_batch(out, 1, &n, box, pos, vel, mass, precision);
}

int pfm_create_batch(pfm_ctx** out, int nrep, const int* n_atoms, const double* boxes, const double* pos,
                     const double* vel, const double* mass, int precision) {
  pfm_ctx* h = new pfm_ctx();
  *out = h;
  if (nrep <= 0) return fail(h, PFM_E_INVALID, "error: a batch needs at least one replica");
  long nl = 0;
  for (int r = 0; r < nrep; ++r) {
    if (n_atoms[r] <= 0) return fail(h, PFM_E_INVALID, "error: particles_num must be positive");
    for (int d = 0; d < 3; ++d)
      if (!(boxes[3 * r + d] > 0)) return fail(h, PFM_E_INVALID, "error: box size must be positive");
    h->rep_start.push_back((int)nl);
    h->rep_n.push_back(n_atoms[r]);
    nl += n_atoms[r];
  }
  if (nl > (1L << 30)) return fail(h, PFM_E_INVALID, "error: too many atoms in one context");
  const int n = (int)nl;
  h->nrep = nrep;
  h->rep_box.assign(boxes, boxes + 3 * nrep);
  if (precision != PFM_FP64 && precision != PFM_MIXED) return fail(h, PFM_E_INVALID, "error: unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(h, PFM_E_NO_DEVICE, "error: no HIP device visible (libpfmds needs an MI355X / gfx950)");
  HIPCHK(h, hipGetDevice(&h->dev));
  hipDeviceProp_t prop;
  HIPCHK(h, hipGetDeviceProperties(&prop, h->dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(h, PFM_E_NO_DEVICE, std::string("error: libpfmds is built for gfx950, device is ") + prop.gcnArchName);
  HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->N = n;
  h->precision = precision;
  for (int d = 0; d < 3; ++d) {
    h->box[d] = boxes[d];
    h->hbox[d] = 0.5 * boxes[d];
  }
  int rc;
  if ((rc = dalloc(h, &h->pos4, n)) || (rc = dalloc(h, &h->vel4, n)) || (rc = dalloc(h, &h->frc4, n)) ||
      (rc = dalloc(h, &h->pos4b, n)) || (rc = dalloc(h, &h->vel4b, n)) || (rc = dalloc(h, &h->frc4b, n)) ||
      (rc = dalloc(h, &h->mask, n)) || (rc = dalloc(h, &h->maskb, n)) || (rc = dalloc(h, &h->orig, n)) ||
      (rc = dalloc(h, &h->origb, n)) || (rc = dalloc(h, &h->perm, n)) || (rc = dalloc(h, &h->atom_rank, n)) ||
      (rc = dalloc(h, &h->scan_tmp, 3 * (size_t)n + 4096)) || (rc = dalloc(h, &h->eb, n)))
    return rc;
  for (int d = 0; d < 3; ++d) (&h->qscale.x)[d] = 4294967296.0 / boxes[d];
  if (precision == PFM_MIXED && (rc = dalloc(h, &h->qpos, n))) return rc;
  h->red_blocks = std::max(1, std::min(1024, (n + RED_THREADS - 1) / RED_THREADS));
  // compensated partial rows: an (s, c) pair per block
  if ((rc = dalloc(h, &h->red_part, (size_t)2 * (PFM_MAXI + 1 + RED_MAXQ) * h->red_blocks))) return rc;
  if ((rc = dalloc(h, &h->red_out, 64))) return rc;
  if ((rc = dalloc(h, &h->derr, 1))) return rc;
  HIPCHK(h, hipHostMalloc((void**)&h->h_red, sizeof(double) * 64, 0));
  HIPCHK(h, hipHostMalloc((void**)&h->herr, sizeof(DevErr), 0));
  std::vector<double4> p(n), v(n);
  std::vector<int> o(n);
  h->hmass.assign(mass, mass + n);
  for (int i = 0; i < n; ++i) {
    p[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
    v[i] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], mass[i]);
    o[i] = i;
  }
  HIPCHK(h, hipMemcpyAsync(h->pos4, p.data(), sizeof(double4) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->vel4, v.data(), sizeof(double4) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->orig, o.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
  if (nrep > 1) {
    std::vector<int> rp(n);
    for (int r = 0; r < nrep; ++r)
      for (int k = 0; k < h->rep_n[r]; ++k) rp[h->rep_start[r] + k] = r;
    if ((rc = dalloc(h, &h->rep, n)) || (rc = dalloc(h, &h->repb, n))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->rep, rp.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFM_OK;
}

int pfm_set_group(pfm_ctx* h, int gid, int n, const int* idx) {
  if (h->sorted_once) return fail(h, PFM_E_INVALID, "error: groups must be set before the first pfm_run");
  if (gid < 1 || gid > PFM_MAXG) return fail(h, PFM_E_INVALID, "error: group number out of range (1..32)");
  std::vector<int>& g = h->grp.idx[gid];
  g.clear();
  std::vector<char> seen(h->N, 0);
  for (int k = 0; k < n; ++k) {
    int a = idx[k] - 1;
    if (a < 0 || a >= h->N) return fail(h, PFM_E_INVALID, "error: group index out of range");
    if (seen[a]) return fail(h, PFM_E_INVALID, "error: atom listed twice in a group (species named twice)");
    seen[a] = 1;
    g.push_back(a);
  }
  h->grp.n[gid] = n;
  h->grp.size[gid] = n;
  return PFM_OK;
}

int pfm_set_roles(pfm_ctx* h, int all_moving, int xyz_moving, int z_moving, int all_atoms, int zmp, int invert_z) {
  int r[4] = {all_moving, xyz_moving, z_moving, all_atoms};
  for (int k = 0; k < 4; ++k)
    if (r[k] < 1 || r[k] > PFM_MAXG) return fail(h, PFM_E_INVALID, "error: role group out of range");
  if (zmp < 1) return fail(h, PFM_E_INVALID, "error: zero_momentum_period must be >= 1");
  for (int k = 0; k < 4; ++k) h->roles[k] = r[k];
  h->zm_period = zmp;
  h->invert_z = invert_z ? 1 : 0;
  return PFM_OK;
}

int pfm_add_interaction(pfm_ctx* h, int kind, const double* params, int np, int nl_n, const int* gn,
                        const int* nmax, const double* rcut, const int* period) {
  static const int need_np[9] = {0, 4, 4, 6, 7, 7, 10, 1, 18};
  static const int need_nl[9] = {0, 2, 1, 3, 3, 1, 1, 2, 1};
  if (h->sorted_once) return fail(h, PFM_E_INVALID, "error: interactions must be added before the first pfm_run");
  if (kind < 1 || kind > 8) return fail(h, PFM_E_UNKNOWN_KIND, "error: unknown interaction name");
  if ((int)h->inter.size() >= PFM_MAXI) return fail(h, PFM_E_INVALID, "error: too many interactions");
  if (np < need_np[kind] || nl_n != need_nl[kind])
    return fail(h, PFM_E_INVALID, "error: wrong parameter or list count for interaction");
  Inter it;
  it.kind = kind;
  it.nl_n = nl_n;
  for (int i = 0; i < np && i < 24; ++i) it.p[i] = params[i];
  for (int k = 0; k < nl_n; ++k) {
    int g1 = gn[2 * k], g2 = gn[2 * k + 1];
    if (g1 < 1 || g1 > PFM_MAXG || g2 < 1 || g2 > PFM_MAXG)
      return fail(h, PFM_E_INVALID, "error: interaction group out of range");
    if (period[k] < 1) return fail(h, PFM_E_INVALID, "error: update_period must be >= 1");
    if (nmax[k] < 1) return fail(h, PFM_E_INVALID, "error: neighb_num_max must be >= 1");
    ListDev& L = it.nl[k];
    L.g1 = g1; L.g2 = g2; L.nmax = nmax[k]; L.rcut = rcut[k]; L.period = period[k];
    L.nrows = h->grp.size[g1];
    h->grp.used[g1] = true;
  }
  if (kind == PFM_LJ1G || kind == PFM_RJL || kind == PFM_TB || kind == PFM_REBOSC) {
    // these loops index group-2 entries with group-1 arrays
    // (LennardJones_1g.f90:83, RosatoGuillopeLegrand.f90:88, TersoffBrenner.f90:56-57,112,
    // REBOsolidcarbon.f90:49,75)
    if (h->grp.idx[it.nl[0].g1] != h->grp.idx[it.nl[0].g2])
      return fail(h, PFM_E_INVALID, "error: lj1g/rjl/tb/rebosc need the same group on both sides of the list");
  }
  if (kind == PFM_REBOSC && it.nl[0].nmax > 32)
    return fail(h, PFM_E_INVALID, "error: rebosc neighb_num_max above 32 is not supported");
  if (kind == PFM_SUR) {
    std::vector<char> in2(h->N, 0);
    for (int x : h->grp.idx[it.nl[0].g2]) in2[x] = 1;
    for (int x : h->grp.idx[it.nl[0].g1])
      if (!in2[x]) return fail(h, PFM_E_INVALID, "error: sur group 1 must be part of group 2");
    if (it.nl[1].g1 != it.nl[0].g2 || it.nl[1].g2 != it.nl[0].g1)
      return fail(h, PFM_E_INVALID, "error: sur list 2 must be the converse of list 1");
    if (!(it.p[0] > 0.0) || it.nl[0].rcut < it.p[0])
      return fail(h, PFM_E_INVALID, "error: sur list r_cut must be at least the descriptor cutoff");
    h->grp.used[it.nl[0].g2] = true;
  }
  if (kind == PFM_TB && it.nl[0].nmax > PFM_TB_MAXN)
    return fail(h, PFM_E_INVALID, "error: tb neighb_num_max above 16 is not supported");
  if (kind == PFM_LJ1G) h->grp.used[it.nl[0].g2] = true;
  if (is_cos(kind)) {
    if (it.nl[2].nmax != 3) return fail(h, PFM_E_GRAPHENE_NN, "error: nl_nn%neighb_num_max/=nnum_nn");
    if (h->grp.idx[it.nl[2].g1] != h->grp.idx[it.nl[0].g1])
      return fail(h, PFM_E_INVALID, "error: ljc/morsec nearest-neighbour rows must be the graphene group of list 1");
    // md_interactions.f90:159-171: the first tb interaction's list feeds nl(3)
    for (size_t j = 0; j < h->inter.size(); ++j)
      if (h->inter[j].kind == PFM_TB || h->inter[j].kind == PFM_REBOSC) { it.nn_from_tb = (int)j; break; }
    if (it.nn_from_tb >= 0) {
      Inter& tb = h->inter[it.nn_from_tb];
      if (h->grp.idx[tb.nl[0].g1] != h->grp.idx[it.nl[0].g1])
        return fail(h, PFM_E_INVALID, "error: tb group and ljc/morsec graphene group differ");
      it.nl[2].period = tb.nl[0].period;
    }
  }
  if (kind == PFM_TB || kind == PFM_REBOSC) {
    for (auto& o : h->inter)
      if (is_cos(o.kind) && o.nn_from_tb < 0)
        // the reference copies the empty tb list at step 0 and stops
        return fail(h, PFM_E_GRAPHENE_NN, "error: not enough gr nearest neibs (tb listed after ljc/morsec)");
  }
  h->inter.push_back(it);
  return PFM_OK;
}

int pfm_add_nhc(pfm_ctx* h, int group, double T, int M, double q1) {
  if ((int)h->nhc.size() >= PFM_MAXNHC || M < 1 || M > 16 || group < 1 || group > PFM_MAXG)
    return fail(h, PFM_E_INVALID, "error: bad nhc definition");
  int L = h->grp.size[group];
  if (L < 1 || q1 < 0.0 || T < 0.0) return fail(h, PFM_E_NHC_PARAMS, "error: wrong nhc parameters");
  for (auto& c : h->nhc) {
    std::vector<int> a = h->grp.idx[c.group], b = h->grp.idx[group];
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    std::vector<int> x;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(x));
    if (!x.empty()) return fail(h, PFM_E_INVALID, "error: thermostat groups overlap");
  }
  Nhc c;
  c.group = group; c.M = M; c.L = L; c.T = T; c.q1 = q1;
  h->nhc.push_back(c);
  h->grp.used[group] = true;
  return PFM_OK;
}

static int run_steps(pfm_ctx* h, int integ, double dt, int s0, int s1, int energy_at_end, double ms_de,
                     int* last_step, int* event) {
  if (integ < PFM_NVT || integ > PFM_NVMS) return fail(h, PFM_E_INVALID, "error: unknown integrator");
  if (h->roles[0] == 0) return fail(h, PFM_E_INVALID, "error: pfm_set_roles was not called");
  if (h->nrep > 1 && integ == PFM_NVT && !h->nhc.empty())
    return fail(h, PFM_E_INVALID, "error: Nose-Hoover chains span replicas; run batches with nve or nvms");
  for (int k = 0; k < 4; ++k) h->grp.used[h->roles[k]] = true;
  int rc;
  if (!h->sorted_once && (rc = finalize(h))) return rc;
  *event = PFM_EV_DONE;
  *last_step = s0;
  IntP P = int_params(h, dt);
  const int N = h->N, nb = (N + 255) / 256;
  uint32_t all_bit = 1u << (h->roles[3] - 1);
  bool all_grows = false;
  for (const auto& c : h->changes) all_grows = all_grows || c.to == h->roles[3];
  bool everyone = h->grp.size[h->roles[3]] == N && !all_grows;
  bool first_writes = false;
  if (everyone && !h->inter.empty()) {
    Inter& f = h->inter[0];
    first_writes = (f.kind == PFM_RJL || f.kind == PFM_LJ1G || f.kind == PFM_LJ || f.kind == PFM_TB) &&
                   f.nl[0].nrows == N;
    // or the shared launches write every atom first
    first_writes = first_writes || multi_first_writes(h) || gm_first_writes(h);
  }
  // check_positions (and invert_z) of the first step of the call
  PFM_LAUNCH(h, "check", k_check, nb, 256, N, h->pos4, h->box[0], h->box[1], h->box[2], h->derr, h->orig, s0,
             (s0 == 0 && h->invert_z) ? 1 : 0, h->vel4, P.zlo, P.zhi, h->rep, h->rdev);
  if ((rc = check_err(h, s1))) return rc;
  if (h->qpos) PFM_LAUNCH(h, "quantize", k_quantize, nb, 256, N, h->pos4, h->qpos, h->qscale, h->rep, h->rdev);
  if (h->need_forces && s0 != 0) {
    // positions set from the host: the first half kick needs their forces
    if ((rc = update_lists(h, s0))) return rc;
    if (!first_writes) PFM_LAUNCH(h, "zero_forces", k_zero_masked, nb, 256, N, h->frc4, h->mask, all_bit);
    if ((rc = launch_forces(h, false, first_writes))) return rc;
  }
  h->need_forces = false;
  const bool nvt = integ == PFM_NVT && !h->nhc.empty();
  bool nhc_pending = false;
  bool kick_pending = false;  // second half kick deferred into the next kick_drift
  const int CONVERGED = -1;
  // nvms blocks: the steps from spec_base on record their energies and
  // states into ring slots instead of reading energies back
  bool spec_on = false;
  int spec_base = 0;
  // one md() step: the body of md_simulation.f90:137-231
  auto body = [&](int s) -> int {
    int rc;
    *last_step = s;
    // change_particle_group_N at the top of the step (md_simulation.f90:116-119)
    if (!h->changes.empty()) apply_changes(h, s);
    bool energy = (s == s1 && energy_at_end) || integ == PFM_NVMS;
    const bool re = rec_e_at(h, s), rx = rec_x_at(h, s);
    // velocities complete at the end of the step: something reads them, or the
    // next step moves another set of atoms
    const bool report = energy || re || rx || counts_change_at(h, s + 1);
    if (s != 0) {
      bool conv = std::fabs(h->pe - h->pe_prev) <= ms_de;
      if (h->nrep > 1 && integ == PFM_NVMS) {
        // a batch stops when every replica meets the criterion
        conv = h->rep_pe.size() == (size_t)h->nrep && h->rep_pe_prev.size() == (size_t)h->nrep;
        for (int r = 0; conv && r < h->nrep; ++r) conv = std::fabs(h->rep_pe[r] - h->rep_pe_prev[r]) <= ms_de;
      }
      if (integ == PFM_NVMS && conv && !spec_on) return CONVERGED;
      IntP Q = P;
      Q.pre_kick = kick_pending ? 1 : 0;
      kick_pending = false;
      Q.invert = h->invert_z;
      Q.zero_bit = first_writes ? 0u : all_bit;
      if (nvt) {
        if (!nhc_pending) nhc_first_half(h, dt);
        Q.scale = 1;
        nhc_pending = false;
      }
      PFM_LAUNCH(h, "kick_drift", k_kick_drift, nb, 256, N, h->pos4, h->vel4, h->frc4, h->mask, Q, h->derr, h->orig,
                 s - h->step_base, h->qpos, h->qscale, h->rep, h->rdev);
    } else if (!first_writes) {
      PFM_LAUNCH(h, "zero_forces", k_zero_masked, nb, 256, N, h->frc4, h->mask, all_bit);
    }
    h->pe_prev = h->pe;
    h->rep_pe_prev = h->rep_pe;
    if ((rc = update_lists(h, s))) return rc;
    if (s % h->zm_period == 0 && (rc = zero_momentum(h))) return rc;
    if ((rc = launch_forces(h, energy || re, first_writes))) return rc;
    if (s != 0) {
      if (nvt) {
        // second half step; on steps with nothing to report, the chain update
        // also takes the first half of the next step (k_nhc_chain)
        MaskQ C = chain_masks(h);
        bool fuse = !report && s < s1;
        ChainArgs A{(NhcDev*)h->nhc_state, h->red_out + 16, dt / 2.0, dt / 4.0, dt / 8.0, fuse ? 1 : 0};
        PFM_LAUNCH(h, "kick2", k_kick2_ke, h->red_blocks, RED_THREADS, N, h->vel4, h->frc4, h->mask, P, C, h->red_part);
        PFM_LAUNCH(h, "nhc", k_nhc_chain, 1, RED_THREADS, C.nq, A, h->red_part, h->red_blocks);
        if (fuse) nhc_pending = true;
        else PFM_LAUNCH(h, "nhc_scale", k_scale, nb, 256, N, h->vel4, h->mask, P);
      } else if (integ != PFM_NVMS && !report && s < s1) {
        // nothing reads the velocities before the next step's first half
        kick_pending = true;
      } else {
        PFM_LAUNCH(h, "kick2", k_kick2, nb, 256, N, h->vel4, h->frc4, h->mask, P, integ == PFM_NVMS ? 1 : 0);
      }
      h->sim_time = h->sim_time + dt;
    }
    if (spec_on) {
      spec_record(h, s - spec_base);
    } else if (energy) {
      if ((rc = energies(h, integ == PFM_NVT))) return rc;
      if ((rc = check_err(h, s))) return rc;
    }
    if (re || rx) rec_launch(h, re, rx);
    return PFM_OK;
  };
  // Steady-state blocks of `blk` steps (the list and zero-momentum schedule
  // repeats with that period) that start on a multiple of blk and hold no
  // energy step are captured once as a HIP graph and replayed: the launches
  // of a block depend only on the schedule, the run's parameters and which
  // of the two re-sort buffers is current, so the cache is keyed by those.
  const int blk = graph_block(h);
  const bool graphs = h->use_graphs && !h->timing && integ != PFM_NVMS && blk > 0;
  // steps s .. s+nblk-1 from a captured graph: captured at their first
  // occurrence (the host side of body runs then), replayed afterwards
  auto run_block = [&](int s, int nblk) -> int {
    int rc;
    std::vector<uint64_t> key = graph_key(h, integ, dt, nblk, first_writes, s);
    key.push_back((uint64_t)(spec_on ? h->spec_e : nullptr));
    pfm_ctx::GraphEntry* g = nullptr;
    for (auto& e : h->graphs)
      if (e.key == key) g = &e;
    if (!g) {
      hipGraph_t graph;
      h->step_base = s;
      HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
      int rc2 = PFM_OK;
      for (int t = s; t < s + nblk && rc2 == PFM_OK; ++t) rc2 = body(t);
      hipError_t ec = hipStreamEndCapture(h->stream, &graph);
      h->step_base = 0;
      if (rc2 != PFM_OK) {
        if (ec == hipSuccess) (void)hipGraphDestroy(graph);
        return rc2 == CONVERGED ? fail(h, PFM_E_INVALID, "error: convergence test inside a captured block") : rc2;
      }
      HIPCHK(h, ec);
      const bool same = integ == PFM_NVMS ? !kick_pending : (nvt ? nhc_pending : kick_pending);
      if (!same) {
        (void)hipGraphDestroy(graph);
        return fail(h, PFM_E_INVALID, "error: captured block does not end in its starting state");
      }
      pfm_ctx::GraphEntry e;
      e.key = key;
      e.after = graph_buffers(h);
      HIPCHK(h, hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      if (h->graphs.size() >= GRAPH_CACHE) {
        // least recently launched block goes
        auto old = std::min_element(h->graphs.begin(), h->graphs.end(),
                                    [](const pfm_ctx::GraphEntry& a, const pfm_ctx::GraphEntry& b) {
                                      return a.used < b.used;
                                    });
        // the evicted block may still be running: drain the stream first
        HIPCHK(h, hipStreamSynchronize(h->stream));
        (void)hipGraphExecDestroy(old->exec);
        h->graphs.erase(old);
      }
      e.used = ++h->graph_clock;
      h->graphs.push_back(e);
      ++h->graph_captures;
      ++h->graph_launches;
      // the capture ran the block's host side; its kernels run now
      PFM_LAUNCH(h, "graph_base", k_set_base, 1, 1, h->derr, s);
      HIPCHK(h, hipGraphLaunch(h->graphs.back().exec, h->stream));
      PFM_LAUNCH(h, "graph_base", k_set_base, 1, 1, h->derr, 0);
    } else {
      g->used = ++h->graph_clock;
      PFM_LAUNCH(h, "graph_base", k_set_base, 1, 1, h->derr, s);
      HIPCHK(h, hipGraphLaunch(g->exec, h->stream));
      PFM_LAUNCH(h, "graph_base", k_set_base, 1, 1, h->derr, 0);
      ++h->graph_launches;
      // the host side of the block: buffers, time, energies carried over
      set_graph_buffers(h, g->after);
      *last_step = s + nblk - 1;
      for (int t = 0; t < nblk; ++t) h->sim_time = h->sim_time + dt;
      h->pe_prev = h->pe;
      h->rep_pe_prev = h->rep_pe;
    }
    (void)rc;
    return PFM_OK;
  };
  std::vector<RecNote> notes;
  // nvms without a read-back per step: blocks of K steps record every step's
  // energies and state; one read-back per block finds the first step whose
  // exit test holds (md_simulation.f90:142-147) and restores the state of the
  // step before it
  auto nvms_blocks = [&]() -> int {
    int rc;
    const int K = (blk > 0 && blk <= 32) ? blk : 16;
    if ((rc = spec_alloc(h, K))) return rc;
    const int ni = (int)h->inter.size(), w = ni + 1;
    const int nr = h->nrep > 1 ? h->nrep : 0;
    std::vector<double> ev((size_t)K * w * (1 + nr));
    const bool capture = h->use_graphs && blk == K;
    // the exit test of md_simulation.f90:143 between two steps' energies; a
    // batch stops when every replica meets it
    auto conv = [&](double a, double b, const std::vector<double>& ra, const std::vector<double>& rb) {
      if (!nr) return std::fabs(a - b) <= ms_de;
      if ((int)ra.size() != nr || (int)rb.size() != nr) return false;
      for (int r = 0; r < nr; ++r)
        if (!(std::fabs(ra[r] - rb[r]) <= ms_de)) return false;
      return true;
    };
    for (int s = s0; s <= s1;) {
      *last_step = s;
      if (s != 0 && conv(h->pe, h->pe_prev, h->rep_pe, h->rep_pe_prev)) {
        *event = PFM_EV_CONVERGED;
        return exit_step_top(h, s);
      }
      int end = std::min(s1, (s / K + 1) * K - 1);
      int ne_b = 0, nx_b = 0;
      if (rec_on(h)) {
        // the block's records fit the rings: none is handed on before its block is settled
        for (int t = s; t <= end; ++t) {
          int e2 = ne_b + (rec_e_at(h, t) ? 1 : 0), x2 = nx_b + (rec_x_at(h, t) ? 1 : 0);
          if (t > s && (e2 > h->rec.ecap || x2 > h->rec.xcap)) {
            end = t - 1;
            break;
          }
          ne_b = e2;
          nx_b = x2;
        }
        if (ne_b + nx_b > 0 && (rc = rec_reserve(h, ne_b, nx_b))) return rc;
      }
      const int n = end - s + 1;
      const double t_start = h->sim_time, pe0 = h->pe;
      const std::vector<double> rpe0 = h->rep_pe;
      spec_on = true;
      spec_base = s;
      int rc2 = PFM_OK;
      if (capture && s > 0 && s % K == 0 && n == K) {
        rc2 = run_block(s, K);
      } else {
        for (int t = s; t <= end && rc2 == PFM_OK; ++t) rc2 = body(t);
      }
      spec_on = false;
      if (rc2) return rc2;
      if (ne_b + nx_b > 0) {
        notes.clear();
        double tt = t_start;
        for (int t = s; t <= end; ++t) {
          if (t != 0) tt = tt + dt;
          if (rec_e_at(h, t) || rec_x_at(h, t)) notes.push_back(RecNote{t, tt, rec_e_at(h, t), rec_x_at(h, t)});
        }
        if ((rc = rec_issue(h, integ, notes))) return rc;
      }
      HIPCHK(h, hipMemcpyAsync(ev.data(), h->spec_e, sizeof(double) * n * w, hipMemcpyDeviceToHost, h->stream));
      if (nr)
        HIPCHK(h, hipMemcpyAsync(ev.data() + (size_t)K * w, h->spec_e + (size_t)h->spec_k * w,
                                 sizeof(double) * n * nr * w, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipMemcpyAsync(h->herr, h->derr, sizeof(DevErr), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      // pe of every step of the block, summed as energies() sums it; the exit
      // test of step t reads pe(t-1) and pe(t-2)
      std::vector<double> pe(n);
      std::vector<std::vector<double>> rpe(n);
      const double* evr = ev.data() + (size_t)K * w;  // [n][nr][w]
      int tc = -1;
      for (int j = 0; j < n; ++j) {
        double p = 0.0;
        for (int k = 0; k < ni; ++k) p += ev[(size_t)j * w + k];
        pe[j] = p;
        rpe[j].assign(nr, 0.0);
        for (int r = 0; r < nr; ++r)
          for (int k = 0; k < ni; ++k) rpe[j][r] += evr[((size_t)j * nr + r) * w + k];
        const double m2 = j == 0 ? pe0 : pe[j - 1];
        if (j + 1 < n && conv(p, m2, rpe[j], j == 0 ? rpe0 : rpe[j - 1])) {
          tc = s + j + 1;
          break;
        }
      }
      const int jl = tc >= 0 ? tc - 1 - s : n - 1;  // the last step that stands
      for (int k = 0; k < ni; ++k) h->inter[k].energy = ev[(size_t)jl * w + k];
      h->pe = pe[jl];
      h->pe_prev = jl == 0 ? pe0 : pe[jl - 1];
      h->ke = ev[(size_t)jl * w + ni];
      h->temp = 2 * h->ke / PFM_KB / (3 * h->grp.n[h->roles[0]]);
      if (nr) {
        set_rep_energies(h, evr + (size_t)jl * nr * w);
        h->rep_pe_prev = jl == 0 ? rpe0 : rpe[jl - 1];
      }
      if (tc < 0) {
        if ((rc = check_err(h, end)) || (rc = rec_poll(h, false))) return rc;
        s = end + 1;
        continue;
      }
      // steps tc..end ran ahead: their errors and records do not stand
      const DevErr er = *h->herr;
      if (er.code != 0 && er.step < tc) return check_err(h, end);
      if (er.code != 0) HIPCHK(h, hipMemsetAsync(h->derr, 0, sizeof(DevErr), h->stream));
      auto& R = h->rec;
      while (!R.q.empty() && R.q.back().step >= tc) {
        if (R.q.back().eslot >= 0) --R.pending_e;
        if (R.q.back().xslot >= 0) --R.pending_x;
        R.free_ev.push_back(R.q.back().done);
        R.q.pop_back();
      }
      const size_t o = (size_t)jl * h->N;
      PFM_LAUNCH(h, "spec", k_spec_copy, (h->N + 255) / 256, 256, h->N, 1, h->pos4, h->vel4, h->frc4, h->mask,
                 h->orig, h->sp_pos + o, h->sp_vel + o, h->sp_frc + o, h->sp_mask + o, h->sp_orig + o);
      // lists, cells and rows follow the restored state at the next step
      h->need_sort = true;
      h->sim_time = t_start;
      for (int t = s; t < tc; ++t)
        if (t != 0) h->sim_time = h->sim_time + dt;
      *last_step = tc;
      *event = PFM_EV_CONVERGED;
      return exit_step_top(h, tc);
    }
    return check_err(h, s1);
  };
  if (integ == PFM_NVMS && h->speculate && h->changes.empty() && !h->timing) return nvms_blocks();
  for (int s = s0; s <= s1;) {
    const bool steady = nvt ? nhc_pending : kick_pending;
    bool blocked = graphs && s > 0 && s % blk == 0 && s + blk - 1 < s1 && steady;
    if (blocked && !h->changes.empty()) {
      // a block holds no change of group counts, nor does the step after it
      int n[PFM_MAXG + 1];
      std::copy(h->grp.n, h->grp.n + PFM_MAXG + 1, n);
      for (int t = s; t <= s + blk && blocked; ++t) {
        change_counts(n, h->grp.size, h->changes, t);
        blocked = std::equal(n, n + PFM_MAXG + 1, h->grp.n);
      }
    }
    int ne_b = 0, nx_b = 0;
    if (blocked && rec_on(h)) {
      // records inside a block are captured with it; the block must end in
      // the steady state (no record on its last step) and fit the rings
      for (int t = s; t < s + blk; ++t) {
        ne_b += rec_e_at(h, t) ? 1 : 0;
        nx_b += rec_x_at(h, t) ? 1 : 0;
      }
      blocked = !rec_e_at(h, s + blk - 1) && !rec_x_at(h, s + blk - 1) && ne_b <= h->rec.ecap &&
                nx_b <= h->rec.xcap;
    }
    if (blocked) {
      if (ne_b + nx_b > 0 && (rc = rec_reserve(h, ne_b, nx_b))) return rc;
      const double t_start = h->sim_time;
      if ((rc = run_block(s, blk))) return rc;
      if (ne_b + nx_b > 0) {
        notes.clear();
        double tt = t_start;
        for (int t = s; t < s + blk; ++t) {
          tt = tt + dt;
          if (rec_e_at(h, t) || rec_x_at(h, t)) notes.push_back(RecNote{t, tt, rec_e_at(h, t), rec_x_at(h, t)});
        }
        if ((rc = rec_issue(h, integ, notes)) || (rc = rec_poll(h, false))) return rc;
      }
      s += blk;
      continue;
    }
    const bool re = rec_e_at(h, s), rx = rec_x_at(h, s);
    if ((re || rx) && (rc = rec_reserve(h, re ? 1 : 0, rx ? 1 : 0))) return rc;
    int rc2 = body(s);
    if (rc2 == CONVERGED) {
      *event = PFM_EV_CONVERGED;
      return exit_step_top(h, s);
    }
    if (rc2) return rc2;
    if (re || rx) {
      notes.assign(1, RecNote{s, h->sim_time, re, rx});
      if ((rc = rec_issue(h, integ, notes)) || (rc = rec_poll(h, false))) return rc;
    }
    ++s;
  }
  return check_err(h, s1);
}

int pfm_run(pfm_ctx* h, int integ, double dt, int s0, int s1, int energy_at_end, double ms_de, int* last_step,
            int* event) {
  auto& R = h->rec;
  int rc;
  if (rec_on(h)) {
    if (h->nrep > 1) return fail(h, PFM_E_INVALID, "error: records are taken for single systems only");
    if ((rc = rec_prepare(h))) return rc;
    HIPCHK(h, hipMemsetAsync(R.d_ctr, 0, 4 * sizeof(int), h->stream));
    R.ne = R.nx = 0;
  }
  rc = run_steps(h, integ, dt, s0, s1, energy_at_end, ms_de, last_step, event);
  if (rc == PFM_OK && !R.q.empty()) rc = rec_poll(h, true);
  if (rc != PFM_OK && !R.q.empty()) {
    // records of a failed run are dropped
    (void)hipStreamSynchronize(h->stream);
    for (auto& p : R.q) R.free_ev.push_back(p.done);
    R.q.clear();
    R.pending_e = R.pending_x = 0;
  }
  return rc;
}

int pfm_set_records(pfm_ctx* h, int n_e, const int* e_periods, int n_x, const int* x_periods, int frame_slots,
                    pfm_record_fn fn, void* user) {
  auto& R = h->rec;
  if (!R.q.empty()) return fail(h, PFM_E_INVALID, "error: records still pending");
  if (n_e < 0 || n_x < 0 || (n_e > 0 && !e_periods) || (n_x > 0 && !x_periods))
    return fail(h, PFM_E_INVALID, "error: bad record periods");
  std::vector<int> ep, xp;
  for (int k = 0; k < n_e; ++k)
    if (e_periods[k] > 0) ep.push_back(e_periods[k]);
  for (int k = 0; k < n_x; ++k)
    if (x_periods[k] > 0) xp.push_back(x_periods[k]);
  if (!xp.empty() && frame_slots < 1) return fail(h, PFM_E_INVALID, "error: frames need at least one slot");
  if (fn && (!ep.empty() || !xp.empty()) && h->nrep > 1)
    return fail(h, PFM_E_INVALID, "error: records are taken for single systems only");
  // captured blocks hold the records of their steps
  clear_graphs(h);
  R.eper = ep;
  R.xper = xp;
  R.xslots = xp.empty() ? 0 : frame_slots;
  R.fn = fn;
  R.user = user;
  return PFM_OK;
}

int pfm_add_group_change(pfm_ctx* h, int from, int to, int ts1, int ts2, int frec) {
  if (h->sorted_once) return fail(h, PFM_E_INVALID, "error: group changes must be added before the first pfm_run");
  if (from < 1 || from > PFM_MAXG || to < 1 || to > PFM_MAXG)
    return fail(h, PFM_E_INVALID, "error: group number out of range (1..32)");
  if (frec < 1) return fail(h, PFM_E_INVALID, "error: change frequency must be >= 1");
  h->changes.push_back(pfm_ctx::GroupChange{from, to, ts1, ts2, frec});
  return PFM_OK;
}

int pfm_pe_change(pfm_ctx* h, double* de) {
  *de = h->pe - h->pe_prev;
  return PFM_OK;
}

int pfm_set_speculation(pfm_ctx* h, int on) {
  h->speculate = on != 0;
  return PFM_OK;
}

int pfm_group_count(pfm_ctx* h, int gid, int* n) {
  if (gid < 1 || gid > PFM_MAXG) return fail(h, PFM_E_INVALID, "error: group number out of range (1..32)");
  *n = h->grp.n[gid];
  return PFM_OK;
}

int pfm_replica_energies(pfm_ctx* h, double* e_inter, double* pe, double* ke, double* temp) {
  const int nrep = h->nrep, ni = (int)h->inter.size();
  if (nrep == 1) {
    if (e_inter)
      for (int k = 0; k < ni; ++k) e_inter[k] = h->inter[k].energy;
    if (pe) *pe = h->pe;
    if (ke) *ke = h->ke;
    if (temp) *temp = h->temp;
    return PFM_OK;
  }
  if (h->rep_pe.size() != (size_t)nrep) return fail(h, PFM_E_INVALID, "error: no energies evaluated yet");
  for (int r = 0; r < nrep; ++r) {
    if (e_inter)
      for (int k = 0; k < ni; ++k) e_inter[(size_t)r * ni + k] = h->rep_e[(size_t)r * (ni + 1) + k];
    if (pe) pe[r] = h->rep_pe[r];
    if (ke) ke[r] = h->rep_ke[r];
    if (temp) temp[r] = h->rep_temp[r];
  }
  return PFM_OK;
}

int pfm_energies(pfm_ctx* h, double* e_inter, double* pe, double* ke, double* temp, double* e_nhc, double* t) {
  if (e_inter)
    for (size_t i = 0; i < h->inter.size(); ++i) e_inter[i] = h->inter[i].energy;
  if (pe) *pe = h->pe;
  if (ke) *ke = h->ke;
  if (temp) *temp = h->temp;
  if (e_nhc)
    for (size_t i = 0; i < h->e_nhc.size(); ++i) e_nhc[i] = h->e_nhc[i];
  if (t) *t = h->sim_time;
  return PFM_OK;
}

static int download_orig(pfm_ctx* h, std::vector<int>& o) {
  o.resize(h->N);
  HIPCHK(h, hipMemcpyAsync(o.data(), h->orig, sizeof(int) * h->N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFM_OK;
}

int pfm_get_state(pfm_ctx* h, double* pos, double* vel, double* force) {
  std::vector<int> o;
  int rc;
  if ((rc = download_orig(h, o))) return rc;
  std::vector<double4> buf(h->N);
  double4* src[3] = {h->pos4, h->vel4, h->frc4};
  double* dst[3] = {pos, vel, force};
  for (int k = 0; k < 3; ++k) {
    if (!dst[k]) continue;
    HIPCHK(h, hipMemcpyAsync(buf.data(), src[k], sizeof(double4) * h->N, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < h->N; ++i) {
      int a = o[i];
      dst[k][3 * a] = buf[i].x; dst[k][3 * a + 1] = buf[i].y; dst[k][3 * a + 2] = buf[i].z;
    }
  }
  return PFM_OK;
}

int pfm_set_state(pfm_ctx* h, const double* pos, const double* vel) {
  std::vector<int> o;
  int rc;
  if ((rc = download_orig(h, o))) return rc;
  std::vector<double4> p(h->N), v(h->N);
  HIPCHK(h, hipMemcpyAsync(p.data(), h->pos4, sizeof(double4) * h->N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(v.data(), h->vel4, sizeof(double4) * h->N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int i = 0; i < h->N; ++i) {
    int a = o[i];
    if (pos) { p[i].x = pos[3 * a]; p[i].y = pos[3 * a + 1]; p[i].z = pos[3 * a + 2]; }
    if (vel) { v[i].x = vel[3 * a]; v[i].y = vel[3 * a + 1]; v[i].z = vel[3 * a + 2]; }
  }
  HIPCHK(h, hipMemcpyAsync(h->pos4, p.data(), sizeof(double4) * h->N, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->vel4, v.data(), sizeof(double4) * h->N, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto& it : h->inter)
    for (int k = 0; k < it.nl_n; ++k) it.nl[k].built = false;
  // the cells of the last sort do not hold the new positions, nor do the forces
  if (h->sorted_once) h->need_sort = true;
  h->need_forces = true;
  return PFM_OK;
}

// atom data into original order: out[orig[i]] = in[i] (two arrays at once)
__global__ void k_to_orig(int N, const int* __restrict__ orig, const double4* __restrict__ a,
                          const double4* __restrict__ b, double4* __restrict__ oa, double4* __restrict__ ob) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int o = orig[i];
  oa[o] = a[i];
  ob[o] = b[i];
}

// positions and velocities (masses kept) from original order into the handle's order
__global__ void k_from_orig(int N, const int* __restrict__ orig, const double4* __restrict__ pa,
                            const double4* __restrict__ va, double4* __restrict__ pos, double4* __restrict__ vel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int o = orig[i];
  const double4 p = pa[o], v = va[o];
  double4 q = pos[i], w = vel[i];
  q.x = p.x; q.y = p.y; q.z = p.z;
  w.x = v.x; w.y = v.y; w.z = v.z;
  pos[i] = q;
  vel[i] = w;
}

// per-block partial sums of |fa - fb|^2 over the atoms of the group bit in
// handle b's order, fa given in original order
__global__ void __launch_bounds__(RED_THREADS) k_force_dev(int N, const int* __restrict__ orig,
                                                           const double4* __restrict__ fa_orig,
                                                           const double4* __restrict__ fb,
                                                           const uint32_t* __restrict__ mask, uint32_t bit,
                                                           double* __restrict__ part) {
  __shared__ double sh[2 * RED_THREADS];
  CSum a{0.0, 0.0};
  for (int i = blockIdx.x * RED_THREADS + threadIdx.x; i < N; i += gridDim.x * RED_THREADS) {
    if (!(mask[i] & bit)) continue;
    const double4 x = fa_orig[orig[i]], y = fb[i];
    const double dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z;
    a.add(dx * dx + dy * dy + dz * dz);
  }
  CSum s = block_csum(a, sh);
  if (threadIdx.x == 0) put_part(part, 0, gridDim.x, blockIdx.x, s);
}

// the staging of handle h, written next on another handle's stream: h's own
// stream (its clearing of the new buffer, earlier readers) is finished first
static int xfer_ready(pfm_ctx* h) {
  int rc;
  if (!h->xfer && (rc = dalloc(h, &h->xfer, 2 * (size_t)h->N))) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFM_OK;
}

int pfm_copy_state(pfm_ctx* dst, pfm_ctx* src) {
  if (dst->N != src->N || dst->dev != src->dev)
    return fail(dst, PFM_E_INVALID, "error: state copies need handles of one size on one device");
  int rc;
  if ((rc = xfer_ready(dst))) return rc;
  const int N = src->N, nb = (N + 255) / 256;
  // src's positions and velocities in original order, written on src's stream
  PFM_LAUNCH(src, "xfer", k_to_orig, nb, 256, N, src->orig, src->pos4, src->vel4, dst->xfer, dst->xfer + N);
  HIPCHK(src, hipStreamSynchronize(src->stream));
  PFM_LAUNCH(dst, "xfer", k_from_orig, nb, 256, N, dst->orig, dst->xfer, dst->xfer + N, dst->pos4, dst->vel4);
  for (auto& it : dst->inter)
    for (int k = 0; k < it.nl_n; ++k) it.nl[k].built = false;
  if (dst->sorted_once) dst->need_sort = true;
  dst->need_forces = true;
  return PFM_OK;
}

int pfm_force_rmse(pfm_ctx* a, pfm_ctx* b, int group, double* rmse) {
  if (a->N != b->N || a->dev != b->dev)
    return fail(b, PFM_E_INVALID, "error: force comparisons need handles of one size on one device");
  if (group < 1 || group > PFM_MAXG || b->grp.n[group] == 0)
    return fail(b, PFM_E_INVALID, "error: force comparison over an empty or unknown group");
  int rc;
  if ((rc = xfer_ready(b))) return rc;
  const int N = a->N, nb = (N + 255) / 256;
  PFM_LAUNCH(a, "xfer", k_to_orig, nb, 256, N, a->orig, a->frc4, a->frc4, b->xfer, b->xfer + N);
  HIPCHK(a, hipStreamSynchronize(a->stream));
  PFM_LAUNCH(b, "xfer", k_force_dev, b->red_blocks, RED_THREADS, N, b->orig, b->xfer, b->frc4, b->mask,
             1u << (group - 1), b->red_part);
  PFM_LAUNCH(b, "xfer", k_reduce_rows, 1, RED_THREADS, 1, b->red_blocks, b->red_part, b->red_out + 60);
  HIPCHK(b, hipMemcpyAsync(b->h_red + 60, b->red_out + 60, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(b, hipStreamSynchronize(b->stream));
  *rmse = std::sqrt(b->h_red[60] / (3.0 * b->grp.n[group]));
  return PFM_OK;
}

// ------------------------------------------- force RMSE without allocation
//
// pfm_force_rmse_noalloc runs the comparison of pfm_force_rmse with its
// staging (the forces of handle a in original order) and its partial sums in the
// file-scope device array `scratch` instead of memory the library allocates,
// and calls no allocator, for callers that cannot allocate (code under HIP
// graph capture, serving loops).  Every handle on a device shares that
// device's copy of the array, so calls that overlap on one device race.
#define PFM_NOALLOC_MAX_ATOMS (1 << 20)  // 32 MiB of staging
#define PFM_NOALLOC_PARTS 1024           // blocks of partial sums (red_blocks never exceeds it)
static __device__ __align__(256) double4 scratch[PFM_NOALLOC_MAX_ATOMS + 2 * PFM_NOALLOC_PARTS / 4 + 1];

#define PFM_MAX_DEVICES 64
static std::mutex scratch_mu;
static void* scratch_dev[PFM_MAX_DEVICES];  // address of each device's copy, set on first use

// the device address of `scratch` on h's device, looked up once per device
// with that device current
static int scratch_on(pfm_ctx* h, double4** out) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (h->dev < 0 || h->dev >= PFM_MAX_DEVICES) return fail(h, PFM_E_INVALID, "error: device index out of range");
  if (!scratch_dev[h->dev]) {
    int cur = 0;
    HIPCHK(h, hipGetDevice(&cur));
    if (cur != h->dev) HIPCHK(h, hipSetDevice(h->dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != h->dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(h, PFM_E_HIP, std::string("error: HIP ") + hipGetErrorString(e) + " at hipGetSymbolAddress(scratch)");
    scratch_dev[h->dev] = p;
  }
  *out = (double4*)scratch_dev[h->dev];
  return PFM_OK;
}

// forces into original order: out[orig[i]] = f[i]
__global__ void k_frc_to_orig(int N, const int* __restrict__ orig, const double4* __restrict__ f,
                              double4* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[orig[i]] = f[i];
}

int pfm_force_rmse_noalloc(pfm_ctx* a, pfm_ctx* b, int group, double* rmse) {
  if (a->N != b->N || a->dev != b->dev)
    return fail(b, PFM_E_INVALID, "error: force comparisons need handles of one size on one device");
  if (group < 1 || group > PFM_MAXG || b->grp.n[group] == 0)
    return fail(b, PFM_E_INVALID, "error: force comparison over an empty or unknown group");
  if (a->N > PFM_NOALLOC_MAX_ATOMS || b->red_blocks > PFM_NOALLOC_PARTS)
    return fail(b, PFM_E_INVALID,
                "error: pfm_force_rmse_noalloc takes at most 1048576 atoms (its fixed device scratch); use "
                "pfm_force_rmse");
  int rc;
  double4* st = nullptr;
  if ((rc = scratch_on(b, &st))) return rc;
  double* part = (double*)(st + PFM_NOALLOC_MAX_ATOMS);
  double* out = part + 2 * PFM_NOALLOC_PARTS;  // after the (s, c) partial rows
  const int N = a->N, nb = (N + 255) / 256;
  // a's forces into original order on a's stream (every slot < N is written:
  // orig is a permutation), then the comparison on b's stream
  hipLaunchKernelGGL(k_frc_to_orig, dim3(nb), dim3(256), 0, a->stream, N, a->orig, a->frc4, st);
  HIPCHK(a, hipGetLastError());
  HIPCHK(a, hipStreamSynchronize(a->stream));
  hipLaunchKernelGGL(k_force_dev, dim3(b->red_blocks), dim3(RED_THREADS), 0, b->stream, N, b->orig,
                     (const double4*)st, b->frc4, b->mask, 1u << (group - 1), part);
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(RED_THREADS), 0, b->stream, 1, b->red_blocks, part, out);
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipMemcpyAsync(b->h_red + 61, out, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(b, hipStreamSynchronize(b->stream));
  *rmse = std::sqrt(b->h_red[61] / (3.0 * b->grp.n[group]));
  return PFM_OK;
}

int pfm_get_chain(pfm_ctx* h, int c, double* x, double* v) {
  if (c < 1 || c > (int)h->nhc.size()) return fail(h, PFM_E_INVALID, "error: chain number out of range");
  if (!h->nhc_state) {
    // before the first run: the state set up by pfm_add_nhc
    for (int i = 0; i < h->nhc[c - 1].M; ++i) { x[i] = 0.0; v[i] = 0.0; }
    return PFM_OK;
  }
  NhcDev d;
  HIPCHK(h, hipMemcpyAsync(&d, (NhcDev*)h->nhc_state + (c - 1), sizeof d, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int i = 0; i < d.M; ++i) { x[i] = d.x[i]; v[i] = d.v[i]; }
  return PFM_OK;
}

int pfm_set_chain(pfm_ctx* h, int c, const double* x, const double* v) {
  if (c < 1 || c > (int)h->nhc.size()) return fail(h, PFM_E_INVALID, "error: chain number out of range");
  if (!h->nhc_state) return fail(h, PFM_E_INVALID, "error: chain states are set after the first pfm_run");
  NhcDev d;
  NhcDev* dst = (NhcDev*)h->nhc_state + (c - 1);
  HIPCHK(h, hipMemcpyAsync(&d, dst, sizeof d, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int i = 0; i < d.M; ++i) { d.x[i] = x[i]; d.v[i] = v[i]; }
  HIPCHK(h, hipMemcpyAsync(dst, &d, sizeof d, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PFM_OK;
}

int pfm_get_list(pfm_ctx* h, int inter, int k, int* n_rows, int* nnum, int* nlist, int* lessnnum, int stride) {
  if (inter < 1 || inter > (int)h->inter.size()) return fail(h, PFM_E_INVALID, "error: bad interaction index");
  Inter& it = h->inter[inter - 1];
  if (k < 1 || k > it.nl_n) return fail(h, PFM_E_INVALID, "error: bad list index");
  ListDev& L = it.nl[k - 1];
  *n_rows = L.nrows;
  if (!nnum) return PFM_OK;
  if (!L.built || L.nrows == 0) {
    for (int r = 0; r < L.nrows; ++r) { nnum[r] = 0; if (lessnnum) lessnnum[r] = 0; }
    return PFM_OK;
  }
  std::vector<int> o;
  int rc;
  if ((rc = download_orig(h, o))) return rc;
  std::vector<int> rows(L.nrows), cnt(L.nrows), nb((size_t)L.nrows * L.stride);
  HIPCHK(h, hipMemcpyAsync(rows.data(), h->grp.rows[L.g1], sizeof(int) * L.nrows, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(cnt.data(), L.nnum, sizeof(int) * L.nrows, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(nb.data(), L.nbr, sizeof(int) * nb.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (it.kind == PFM_LJ1G)  // lj1g entries carry their half-list weights in bits 30-31
    for (int& x : nb) x &= 0x3fffffff;
  std::vector<int> loc1(h->N, -1), loc2(h->N, -1);
  for (size_t q = 0; q < h->grp.idx[L.g1].size(); ++q) loc1[h->grp.idx[L.g1][q]] = (int)q;
  int g2 = (k == 3 && it.nn_from_tb >= 0) ? L.g1 : L.g2;
  for (size_t q = 0; q < h->grp.idx[g2].size(); ++q) loc2[h->grp.idx[g2][q]] = (int)q;
  for (int r = 0; r < L.nrows; ++r) {
    int a = o[rows[r]];
    int row = loc1[a];
    int n = cnt[r];
    if (n > stride) return fail(h, PFM_E_INVALID, "error: stride smaller than a row");
    std::vector<int> e(n);
    for (int p = 0; p < n; ++p) e[p] = loc2[o[nb[(size_t)r * L.stride + p]]];
    std::sort(e.begin(), e.end());
    nnum[row] = n;
    int less = -1;
    for (int p = 0; p < n; ++p) {
      nlist[(size_t)row * stride + p] = e[p] + 1;
      if (less < 0 && h->grp.idx[g2][e[p]] > a) less = p;
    }
    if (lessnnum) lessnnum[row] = less < 0 ? n : less;
  }
  return PFM_OK;
}

int pfm_row_energies(pfm_ctx* h, int inter, int* n_rows, double* e_row) {
  if (inter < 1 || inter > (int)h->inter.size()) return fail(h, PFM_E_INVALID, "error: no such interaction");
  const Inter& it = h->inter[inter - 1];
  *n_rows = it.e_rows;
  if (e_row && it.e_rows > 0) {
    if (!it.e_row) return fail(h, PFM_E_INVALID, "error: no energy evaluated yet");
    HIPCHK(h, hipMemcpyAsync(e_row, it.e_row, sizeof(double) * it.e_rows, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return PFM_OK;
}

int pfm_get_normals(pfm_ctx* h, int inter, double* out) {
  if (inter < 1 || inter > (int)h->inter.size()) return fail(h, PFM_E_INVALID, "error: bad interaction index");
  Inter& it = h->inter[inter - 1];
  if (!is_cos(it.kind) || !it.gnorm) return fail(h, PFM_E_INVALID, "error: interaction has no graphene normals");
  std::vector<int> o;
  int rc;
  if ((rc = download_orig(h, o))) return rc;
  std::vector<double4> g(h->N);
  HIPCHK(h, hipMemcpyAsync(g.data(), it.gnorm, sizeof(double4) * h->N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<int> loc(h->N, -1);
  const std::vector<int>& gi = h->grp.idx[it.nl[2].g1];
  for (size_t q = 0; q < gi.size(); ++q) loc[gi[q]] = (int)q;
  for (int i = 0; i < h->N; ++i) {
    int l = loc[o[i]];
    if (l < 0) continue;
    out[3 * l] = g[i].x; out[3 * l + 1] = g[i].y; out[3 * l + 2] = g[i].z;
  }
  return PFM_OK;
}

int pfm_set_timing(pfm_ctx* h, int on) {
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->timing = on != 0;
  h->timers.clear();
  h->evnext = 0;
  return PFM_OK;
}

int pfm_kernel_time(pfm_ctx* h, const char* name, int* count, double* mean_ms) {
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int c = 0;
  double tot = 0.0;
  for (auto& t : h->timers) {
    if (t.name.find(name) == std::string::npos) continue;
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, t.a, t.b));
    tot += ms;
    ++c;
  }
  *count = c;
  *mean_ms = c ? tot / c : 0.0;
  return PFM_OK;
}

int pfm_set_tuning(pfm_ctx* h, int lanes, int block) {
  if (lanes) {
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32)
      return fail(h, PFM_E_INVALID, "error: lanes_per_row must be 1,2,4,8,16 or 32");
    h->lanes = lanes;
  }
  if (block) {
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
      return fail(h, PFM_E_INVALID, "error: block must be 64, 128, 256, 512 or 1024");
    h->block = block;
  }
  return PFM_OK;
}

int pfm_set_fusion(pfm_ctx* h, int on) {
  h->fuse = on != 0;
  return PFM_OK;
}

int pfm_set_pass_lanes(pfm_ctx* h, int inter, int lanes) {
  if (inter < 1 || inter > (int)h->inter.size()) return fail(h, PFM_E_INVALID, "error: no such interaction");
  if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32)
    return fail(h, PFM_E_INVALID, "error: lanes_per_row must be 0,1,2,4,8,16 or 32");
  h->inter[inter - 1].lanes = lanes;
  return PFM_OK;
}

int pfm_last_error(pfm_ctx* h, char* buf, int len) {
  if (len > 0) {
    std::strncpy(buf, h->err.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return PFM_OK;
}

int pfm_set_graphs(pfm_ctx* h, int on) {
  h->use_graphs = on != 0;
  if (!h->use_graphs) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    clear_graphs(h);
  }
  return PFM_OK;
}

int pfm_graph_count(pfm_ctx* h, long* captured, long* launched) {
  if (captured) *captured = h->graph_captures;
  if (launched) *launched = h->graph_launches;
  return PFM_OK;
}

int pfm_destroy(pfm_ctx* h) {
  if (!h) return PFM_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  clear_graphs(h);
  rec_free(h);
  for (auto e : h->rec.free_ev) (void)hipEventDestroy(e);
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(h->pos4); F(h->vel4); F(h->frc4); F(h->pos4b); F(h->vel4b); F(h->frc4b); F(h->mask); F(h->maskb);
  F(h->orig); F(h->origb); F(h->perm); F(h->atom_rank); F(h->scan_tmp); F(h->eb); F(h->qpos); F(h->rdev); F(h->rep); F(h->repb); F(h->cell_rep); F(h->d_rep_row0); F(h->d_rep_out);
  F(h->red_part); F(h->red_out); F(h->derr); F(h->cell_rank); F(h->cell_count); F(h->cell_start); F(h->cell_xyz);
  F(h->nhc_state);
  F(h->xfer);
  F(h->spec_e); F(h->sp_pos); F(h->sp_vel); F(h->sp_frc); F(h->sp_mask); F(h->sp_orig);
  if (h->h_red) (void)hipHostFree(h->h_red);
  if (h->herr) (void)hipHostFree(h->herr);
  for (int g = 0; g <= PFM_MAXG; ++g) { F(h->grp.rows[g]); F(h->grp.row_of[g]); F(h->grp.loc[g]); F(h->grp.rank[g]); F(h->grp.d_rn[g]); }
  for (auto& it : h->inter) {
    for (int k = 0; k < 3; ++k) { F(it.nl[k].nnum); F(it.nl[k].nbr); F(it.nl[k].rev); }
    F(it.e_row); F(it.lc); F(it.B); F(it.Bp); F(it.gnorm); F(it.gC); F(it.gnf); F(it.mrows);
    sur_free(it);
  }
  for (auto e : h->evpool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PFM_OK;
}

}  // extern "C"
