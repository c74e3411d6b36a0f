// pfm_internal.h -- shared definitions of the gfx950 HIP implementation of the
// PFMDS step.  Device layout (all in HBM, atoms in Morton-sorted cell order):
//   pos4[N]  double4 (x, y, z, -)          vel4[N] double4 (vx, vy, vz, mass)
//   frc4[N]  double4 (fx, fy, fz, -)       mask[N] uint32 group bits (bit g-1)
//   orig[N]  int  original 0-based atom index of the internal slot
// Neighbour lists are full lists of internal atom indices, one padded row per
// group-1 atom (rows in internal order); distances are never stored: every
// pass recomputes dr from pos4.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <deque>
#include <string>
#include <vector>

#include "../../include/pfm.h"

#define PFM_MAXG 32
#define PFM_MAXI 16
#define PFM_MAXNHC 8
#define PFM_TB_MAXN 16  // max Tersoff-Brenner list row held in registers

// pi as written by the reference cutoff module (cut_off_function.f90:8,20)
#define PFM_PI 3.14159265358979
// unit constants, left to right as md_general.f90:165,304
#define PFM_MASS_COEF (1.6605389217 / 1.6021765654 * 100.0)
#define PFM_KB (1.3806488 / 1.6021765654 * (1.0 / 10000.0))

// ----------------------------------------------------------------- device math

// find_distance (md_general.f90:423-441), evaluated without contraction so the
// neighbour test dr2 < r_cut**2 is bit-identical to the reference arithmetic.
__device__ __forceinline__ double min_image_dist2(const double4& a, const double4& b, const double3& hb,
                                                  double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
  double x = b.x - a.x, y = b.y - a.y, z = b.z - a.z;
  x = x - hb.x * (copysign(1.0, x - hb.x) + copysign(1.0, x + hb.x));
  y = y - hb.y * (copysign(1.0, y - hb.y) + copysign(1.0, y + hb.y));
  z = z - hb.z * (copysign(1.0, z - hb.z) + copysign(1.0, z + hb.z));
  dx = x; dy = y; dz = z;
  return x * x + y * y + z * z;
}

// Minimum image for the force passes: the same image as find_distance except
// at an exact tie |dx| = L/2 (where only the sign of dx can differ), with
// fused multiply-adds.  Membership tests use min_image_dist2.
__device__ __forceinline__ double fast_dist2(const double4& a, const double4& b, const double3& L, const double3& iL,
                                             double& dx, double& dy, double& dz) {
  double x = b.x - a.x, y = b.y - a.y, z = b.z - a.z;
  x = fma(-L.x, rint(x * iL.x), x);
  y = fma(-L.y, rint(y * iL.y), y);
  z = fma(-L.z, rint(z * iL.z), z);
  dx = x; dy = y; dz = z;
  return fma(x, x, fma(y, y, z * z));
}

// Fixed-point positions of the mixed-precision pair passes: per axis
// u = round(x * 2^32 / L) mod 2^32.  The wrapped int32 difference of two
// atoms' u is their minimum-image separation in units of L / 2^32
// (2.7e-8 A for L = 115 A), exact up to the two quantisation errors, and is
// rounded once to FP32 when scaled back.  Atoms exactly half a box apart
// (never within a cutoff) take the -L/2 image.
__device__ __forceinline__ uint32_t fix32(double x, double s) {
  return (uint32_t)(unsigned long long)__double2ll_rn(x * s);
}
__device__ __forceinline__ uint4 quantize(const double4& p, const double3& s, uint32_t w) {
  return make_uint4(fix32(p.x, s.x), fix32(p.y, s.y), fix32(p.z, s.z), w);
}

// Positions as a pair pass reads them.  Geo<double>: pos4 and the rint
// minimum image in FP64.  Geo<float>: qpos, dr from the fixed-point
// difference.  w() is the per-atom value the previous pass stored beside the
// coordinates.
template <class R> struct Geo;
template <> struct Geo<double> {
  using P = double4;
  const double4* p;
  double3 L, iL;
  __device__ __forceinline__ P at(int j) const { return p[j]; }
  __device__ __forceinline__ double d2(const P& a, const P& b, double& dx, double& dy, double& dz) const {
    return fast_dist2(a, b, L, iL, dx, dy, dz);
  }
  static __device__ __forceinline__ double w(const P& a) { return a.w; }
};
template <> struct Geo<float> {
  using P = uint4;
  const uint4* p;
  float hx, hy, hz;  // L / 2^32 per axis
  __device__ __forceinline__ P at(int j) const { return p[j]; }
  __device__ __forceinline__ float d2(const P& a, const P& b, float& dx, float& dy, float& dz) const {
    dx = (float)(int)(b.x - a.x) * hx;
    dy = (float)(int)(b.y - a.y) * hy;
    dz = (float)(int)(b.z - a.z) * hz;
    return dx * dx + dy * dy + dz * dz;
  }
  static __device__ __forceinline__ float w(const P& a) { return __uint_as_float(a.w); }
};

// Precision policy of the pair arithmetic.  R = double: every operation in
// FP64 with the library's correctly rounded elementary functions.  R = float:
// dr comes from the fixed-point positions (Geo<float>), the pair terms use the gfx950
// transcendental unit (v_exp_f32, v_sin/cos_f32, v_sqrt_f32, v_rcp_f32).
template <class R> struct PM;
template <> struct PM<double> {
  // exponential exp(k*x) with a constant k prepared by kexp(k)
  static __host__ __device__ __forceinline__ double kexp(double k) { return k; }
  static __device__ __forceinline__ double expk(double kx) { return exp(kx); }
  // sin(pi*x), cos(pi*x) with the reference's pi literal
  static __device__ __forceinline__ void sincospi_(double x, double& s, double& c) { sincos(PFM_PI * x, &s, &c); }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double cos_(double x) { return cos(x); }
  static __device__ __forceinline__ double sin_(double x) { return sin(x); }
  static __device__ __forceinline__ double pow_(double x, double y) { return pow(x, y); }
};
template <> struct PM<float> {
  static __host__ __device__ __forceinline__ float kexp(double k) { return (float)(k * 1.4426950408889634); }
  static __device__ __forceinline__ float expk(float kx) { return __builtin_amdgcn_exp2f(kx); }
  // v_sin_f32 / v_cos_f32 take their argument in revolutions
  static __device__ __forceinline__ void sincospi_(float x, float& s, float& c) {
    s = __builtin_amdgcn_sinf(0.5f * x);
    c = __builtin_amdgcn_cosf(0.5f * x);
  }
  static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float cos_(float x) { return __cosf(x); }
  static __device__ __forceinline__ float sin_(float x) { return __sinf(x); }
  static __device__ __forceinline__ float pow_(float x, float y) {
    return x > 0.0f ? __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)) : 0.0f;
  }
};

// cosine cutoff (cut_off_function.f90:6-28): f, and df = f'/r
template <class R>
__device__ __forceinline__ void cos_cutoff(R r, R R1, R R2, R inv_w, R& f, R& df) {
  if (r < R1) {
    f = R(1); df = R(0);
  } else if (r < R2) {
    R arg = R(PFM_PI) * (r - R1) * inv_w;
    f = (R(1) + PM<R>::cos_(arg)) * R(0.5);
    df = -PM<R>::sin_(arg) * R(PFM_PI) * inv_w / r * R(0.5);
  } else {
    f = R(0); df = R(0);
  }
}

// Branch-free cosine cutoff: x = clamp((r-R1)/(R2-R1), 0, 1) gives f = 1,
// df = 0 exactly below R1 and the reference's cosine in between; entries at
// or beyond R2 are masked by the caller.  Returns f and dfr = f'(r)
// (= df_cut * r of the reference, whose df_cut is f'/r), kdf = -pi/(2 (R2-R1)).
template <class R>
__device__ __forceinline__ void cutoff_bl(R r, R R1, R inv_w, R kdf, R& f, R& dfr) {
  R x = (r - R1) * inv_w;
  x = x < R(0) ? R(0) : (x > R(1) ? R(1) : x);
  R sn, cs;
  PM<R>::sincospi_(x, sn, cs);
  f = R(0.5) + R(0.5) * cs;
  dfr = kdf * sn;
}
// Mixed path: the argument in revolutions t = (r - R1) / (2 (R2 - R1)) is one
// fused multiply-add (its constants are loop invariants), clamped to [0, 1/2]
// by one v_med3_f32, then v_sin_f32 / v_cos_f32 of 2 pi t = pi x.
template <>
__device__ __forceinline__ void cutoff_bl<float>(float r, float R1, float inv_w, float kdf, float& f, float& dfr) {
  const float a = 0.5f * inv_w, b = -0.5f * R1 * inv_w;
  const float t = __builtin_amdgcn_fmed3f(__builtin_fmaf(r, a, b), 0.0f, 0.5f);
  f = 0.5f + 0.5f * __builtin_amdgcn_cosf(t);
  dfr = kdf * __builtin_amdgcn_sinf(t);
}

// r = |dr| and 1/r from dr.dr: FP64 with the correctly rounded square root
// and division; the mixed path from one v_rsq_f32 (r = d2 / sqrt(d2))
template <class R>
__device__ __forceinline__ void dist_and_inverse(R d2, R& r, R& ir) {
  r = PM<R>::sqrt_(d2);
  ir = PM<R>::rcp(r);
}
template <>
__device__ __forceinline__ void dist_and_inverse<float>(float d2, float& r, float& ir) {
  ir = __builtin_amdgcn_rsqf(d2);
  r = d2 * ir;
}

// Neumaier's compensated sum (Neumaier 1974, "Rundungsfehleranalyse einiger
// Verfahren zur Summation endlicher Summen"): s holds the running sum, c the
// rounding errors of its additions, and the sum is s + c.  The energy,
// kinetic-energy, momentum and force-deviation reductions carry (s, c) pairs
// from each thread through the block trees and the per-block partials.
struct CSum {
  double s, c;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void add(const CSum& o) {
    add(o.s);
    c += o.c;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// sum over the aligned group of T lanes that share one atom row
template <int T, class V>
__device__ __forceinline__ V group_sum(V v) {
#pragma unroll
  for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
  return v;
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; remap so each XCD (own L2) works on one contiguous range of
// Morton-ordered rows.  Returns -1 for padding blocks.
__device__ __forceinline__ int xcd_block(int nblocks_real) {
  int nb = gridDim.x, b = blockIdx.x;
  int per = nb >> 3;  // grid is padded to a multiple of 8
  int lb = (b & 7) * per + (b >> 3);
  return lb < nblocks_real ? lb : -1;
}

// xcd_block for one sub-range of a launch that holds several passes: the
// sub-range starts on a multiple of 8 blocks and spans np (a multiple of 8)
// blocks, lb is the block's index inside it; returns -1 for padding blocks.
__device__ __forceinline__ int xcd_local(int lb, int np, int nblocks_real) {
  int m = (lb & 7) * (np >> 3) + (lb >> 3);
  return m < nblocks_real ? m : -1;
}

// device error record: code, 1-based atom, count, step.  Kernels replayed
// from a captured graph pass their step relative to the block they were
// captured in; `base` holds the first step of the block being replayed (0
// outside graphs).
struct DevErr {
  int code, atom, count, step, base, pad;
  double x, y, z;  // position of the atom of an out-of-cell error
};
__device__ __forceinline__ void raise_err(DevErr* e, int code, int atom, int count, int step) {
  if (atomicCAS(&e->code, 0, code) == 0) {
    e->atom = atom; e->count = count; e->step = step + e->base;
  }
}
// check_positions (md_general.f90:342-364) reports the atom with its position
__device__ __forceinline__ void raise_out_of_cell(DevErr* e, int atom, int step, const double4& p) {
  if (atomicCAS(&e->code, 0, PFM_E_OUT_OF_CELL) == 0) {
    e->atom = atom; e->count = 0; e->step = step + e->base;
    e->x = p.x; e->y = p.y; e->z = p.z;
  }
}

// Geometry of one replica of a batch (a single system is a batch of one):
// box, minimum-image and fixed-point scales, its cell grid and the first
// global cell rank of that grid.
struct RepDev {
  double3 L, iL, hb;    // box, 1/box, box/2
  double3 qs;           // 2^32 / L
  float3 hq;            // L / 2^32
  float hiso;           // hq.x when the box is a cube (hq equal on every axis), else 0
  double3 clen, inv_clen;
  int nc[3];
  int cell_base;        // global rank (and linear index) of the grid's first cell
  int tile;             // grid at least three cells wide in every direction
  double zlo, zhi;      // invert_z_velocities band
};

// ------------------------------------------------------------- integration

struct IntP {
  double ts0, ts1;          // dt, dt/2
  uint32_t bxyz, bz;        // group bits of xyz-moving and z-moving atoms
  double box[3];
  int nchain;
  uint32_t chain_bit[PFM_MAXNHC];
  const double* chain_s;    // per-chain velocity scale (device)
  int scale, invert;
  double zlo, zhi;
  uint32_t zero_bit;  // zero the forces of these atoms after the first kick (zero_forces)
  int pre_kick;       // first apply the previous step's second half kick (nve)
};

__device__ __forceinline__ double wrap1(double x, double L) {
  if (x > L) return x - L;
  if (x < 0.0) return x + L;
  return x;
}
__device__ __forceinline__ bool in_cell(double x, double L) {
  const double tol = 0.0000001;
  return x > (0.0 - tol) && x < (L + tol);
}

// invert_z_velocities (md_general.f90:382-398) on all atoms
__device__ __forceinline__ void invert_v(double4& v, double z, double lo, double hi) {
  double mid = (lo + hi) / 2;
  if ((z > lo && z < mid && v.z > 0.0) || (z < hi && z > mid && v.z < 0.0)) v.z = -v.z;
}

// First half of a step for atom i: [the previous step's second half kick
// (deferred), same operations as k_kick2], [invert_z_velocities], [NHC
// scaling], kick xyz/z, drift xyz/z with single wrap (md_integrators.f90:7-97),
// the fixed-point copy, zero_forces, and check_positions of the next step
// (md_general.f90:342-364).  The position goes to pos_out and the fixed-point
// copy to q_out; with out_of_place both are written for every atom (moving or
// not), otherwise only where the atom moved.
__device__ __forceinline__ void kick_drift_atom(int i, double4 p, uint32_t m, const IntP& P, const int* rep,
                                                const RepDev* rdev, double3 qs, double4* pos_out, double4* vel,
                                                double4* frc, uint4* q_out, bool out_of_place, DevErr* err,
                                                const int* orig, int step) {
  double Lx = P.box[0], Ly = P.box[1], Lz = P.box[2], zlo = P.zlo, zhi = P.zhi;
  if (rep) {
    const RepDev& R = rdev[rep[i]];
    Lx = R.L.x; Ly = R.L.y; Lz = R.L.z; zlo = R.zlo; zhi = R.zhi; qs = R.qs;
  }
  // the force is read whether or not the atom moves, so its load does not
  // wait on the mask's
  double4 v = vel[i];
  const double4 f = frc[i];
  bool bx = (m & P.bxyz) != 0, bz = (m & P.bz) != 0;
  if (P.pre_kick && (bx || bz)) {
    double k = 1.0 / v.w / PFM_MASS_COEF;
    if (bx) { v.x = v.x + f.x * k * P.ts1; v.y = v.y + f.y * k * P.ts1; v.z = v.z + f.z * k * P.ts1; }
    if (bz) v.z = v.z + f.z * k * P.ts1;
  }
  if (P.invert) invert_v(v, p.z, zlo, zhi);
  if (P.scale) {
    for (int c = 0; c < P.nchain; ++c)
      if (m & P.chain_bit[c]) { double s = P.chain_s[c]; v.x *= s; v.y *= s; v.z *= s; }
  }
  if (bx || bz) {
    double k = 1.0 / v.w / PFM_MASS_COEF;
    if (bx) { v.x = v.x + f.x * k * P.ts1; v.y = v.y + f.y * k * P.ts1; v.z = v.z + f.z * k * P.ts1; }
    if (bz) v.z = v.z + f.z * k * P.ts1;
    if (bx) {
      p.x = wrap1(p.x + v.x * P.ts0, Lx);
      p.y = wrap1(p.y + v.y * P.ts0, Ly);
      p.z = wrap1(p.z + v.z * P.ts0, Lz);
    }
    if (bz) p.z = wrap1(p.z + v.z * P.ts0, Lz);
  }
  if (bx || bz || out_of_place) {
    pos_out[i] = p;
    if (q_out) q_out[i] = quantize(p, qs, 0u);
  }
  vel[i] = v;
  if (m & P.zero_bit) frc[i] = make_double4(0, 0, 0, 0);
  if (!(in_cell(p.x, Lx) && in_cell(p.y, Ly) && in_cell(p.z, Lz)))
    raise_out_of_cell(err, orig[i] + 1, step + 1, p);
}

// ------------------------------------------------------------------ host side

struct ListDev {
  int g1 = 0, g2 = 0, nmax = 0, period = 1, stride = 0;
  double rcut = 0;
  int nrows = 0;
  int* nnum = nullptr;  // [nrows]
  int* nbr = nullptr;   // [nrows*stride] internal atom indices
  int* rev = nullptr;   // [nrows*stride] slot of the row atom in the neighbour's row (tb, nn lists)
  bool built = false;
  bool sparse_cells = false;  // the replicas holding its rows average few atoms per cell: one wave per cell
  // Pair search: an lj / lj1g list and another with the same rows, cutoff and
  // rebuild steps, both searched one wave per cell, are built in one pass
  // (launch_build_pair) by the earlier of the two in build order (the lead)
  int pair_ii = -1, pair_k = -1;
  bool pair_lead = false;
  bool pair_sparse = false;  // the union of the two candidate groups is under half of the atoms
  int done_step = -1;        // the step whose update built this list with its lead
  // Quad search (a pair's lead only): a second pair whose rows are another
  // group and whose lists take the same two candidate groups, built in the
  // same pass; b1 takes this lead's candidate group, b2 its partner's
  int quad_b1i = -1, quad_b1k = -1, quad_b2i = -1, quad_b2k = -1;
  int quad_done = -1;  // the step whose update built this list in a quad lead's pass
};

// device buffers of a surrogate interaction (pfm_surrogate.hip)
struct SurBuf {
  float *G = nullptr, *M = nullptr, *gG = nullptr;  // descriptors, moments, dE/dG per row
  float4* d = nullptr;                              // dE_i/d(dr_ij) per list entry
  double4* fself = nullptr;                         // sum_j dE_i/d(dr_ij) per row
  float* fp = nullptr;                              // biases, output weights, feature scaling
  uint16_t* W = nullptr;                            // BF16 W1, W1^T, W2, W2^T
  bool have_params = false;
  // training
  float* dG = nullptr;
  float4* res = nullptr;
  double* fref = nullptr;
  uint16_t* arr = nullptr;
  float *bpart = nullptr, *wpart = nullptr;
  double *rpart = nullptr, *scal = nullptr, *grad = nullptr;
  int Np = 0, S = 0;
};

struct Inter {
  int kind = 0, nl_n = 0, g[6] = {0, 0, 0, 0, 0, 0};
  double p[24] = {0};
  ListDev nl[3];
  int nn_from_tb = -1;  // ljc/morsec: index of the tb interaction whose list feeds nl(3)
  int lanes = 0;        // lanes per row of its separately launched row passes (0: the handle's)
  double energy = 0.0;
  double* e_row = nullptr;  // per-row energies of the last energy step
  int e_rows = 0;
  // lj1g half-list bookkeeping (LennardJones_1g.f90:42,73 via lessnnum)
  int2* lc = nullptr;  // [N] by internal atom: (group-local index, bound of the half list)
  // tb tables
  double* B = nullptr;   // [nrows*stride] bond order B(p,i)
  double* Bp = nullptr;  // [nrows*stride] B**(1/delt+1)
  // ljc/morsec
  double4* gnorm = nullptr;  // [N] by internal atom
  bool rows_disjoint = false;  // groups 1 and 2 share no atom (passes over both may share a launch)
  double4* gC = nullptr;     // [3N] per graphene atom j and nearest neighbour l: c_l n_j of the normal-derivative term
  float4* gnf = nullptr;     // [N] FP32 copy of gnorm for the mixed metal pass
  int* mrows = nullptr;      // metal rows (nl(2)) with entries, in any order, and their count after them
  // surrogate
  SurBuf* sur = nullptr;
  std::vector<double> sur_theta;  // parameters set before the buffers exist
};

struct Nhc {
  int group = 0, M = 1, L = 0;
  double T = 0, q1 = 0;
};

struct KTimer {
  std::string name;
  hipEvent_t a, b;
};

struct Groups {
  int n[PFM_MAXG + 1] = {0};     // active atoms: the first n of idx (change_particle_group_N)
  int size[PFM_MAXG + 1] = {0};  // members
  // a group that grows at run time keeps its members under a spare bit (the
  // list rows) and its active atoms under its own bit
  int shadow[PFM_MAXG + 1] = {0};
  int* rank[PFM_MAXG + 1] = {nullptr};  // device: original atom -> position in idx or -1 (growing groups);
                                        // in a batch the position among its replica's members
  // batches: active and member counts of every group per replica, and the
  // device copy of a growing group's per-replica counts
  std::vector<int> rn[PFM_MAXG + 1], rsize[PFM_MAXG + 1];
  int* d_rn[PFM_MAXG + 1] = {nullptr};
  std::vector<int> idx[PFM_MAXG + 1];  // host: 0-based original indices, group order
  int* rows[PFM_MAXG + 1] = {nullptr};    // device: members in internal order
  int* row_of[PFM_MAXG + 1] = {nullptr};  // device: internal atom -> row or -1
  int* loc[PFM_MAXG + 1] = {nullptr};     // device: original atom -> group-local index or -1
  bool used[PFM_MAXG + 1] = {false};
  bool ident[PFM_MAXG + 1] = {false};  // identity rows written (a group of every atom)
};

struct pfm_ctx {
  int N = 0, precision = PFM_FP64, dev = 0;
  double box[3], hbox[3];  // replica 0 (the only one of a single system)
  // batch of independent replicas: atoms concatenated in original order
  int nrep = 1;
  std::vector<int> rep_n, rep_start;  // atoms and first original atom of every replica
  std::vector<double> rep_box;        // 3 per replica
  std::vector<RepDev> hrep;           // host copy of rdev
  RepDev* rdev = nullptr;             // [nrep]
  int *rep = nullptr, *repb = nullptr;  // [N] replica of every internal atom (batches only)
  int* cell_rep = nullptr;            // [ncell] replica of every cell (batches only)
  bool any_tile = false, any_rows = false;  // some replica searched by cell tiles / by rows
  // per replica and group: first row of the replica in the group's rows
  std::vector<int> rep_row0[PFM_MAXG + 1];
  int* d_rep_row0 = nullptr;  // [(PFM_MAXG + 1) * (nrep + 1)]; slot 0: first internal atom of every replica
  double* d_rep_out = nullptr;  // [nrep * (n_interactions + 1)] per-replica energies
  std::vector<double> rep_pe, rep_pe_prev, rep_ke, rep_temp, rep_e;  // per-replica energies (batches)
  hipStream_t stream = nullptr;
  // atom state (internal order) and the spare buffers of the re-sort
  double4 *pos4 = nullptr, *vel4 = nullptr, *frc4 = nullptr;
  double4 *pos4b = nullptr, *vel4b = nullptr, *frc4b = nullptr;
  uint32_t *mask = nullptr, *maskb = nullptr;
  int *orig = nullptr, *origb = nullptr;
  int* perm = nullptr;
  const int* bin = nullptr;  // cell-ordered atoms of the last binning when the data was not permuted
  uint32_t host_mask_of_group[PFM_MAXG + 1];
  std::vector<uint32_t> hmask;
  std::vector<double> hmass;
  Groups grp;
  int roles[4] = {0, 0, 0, 0};  // all_moving, xyz, z, all_atoms
  int zm_period = 1, invert_z = 0;
  bool all_atoms_is_everyone = true;
  std::vector<Inter> inter;
  std::vector<Nhc> nhc;
  // NHC chain state on device: x[16], v[16], q[16], s, e, ke per chain
  double* nhc_state = nullptr;
  // cells
  int nc[3] = {1, 1, 1}, ncell = 1;
  double cell_len[3];
  int *cell_rank = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr, *atom_rank = nullptr;
  uint32_t* cell_grp = nullptr;  // inside the cell_count allocation
  // the binning's own counts (inside the cell_count allocation): counted by
  // k_cell_keys, copied to cell_count and cleared with cell_fill by
  // k_cell_sort, so they start every binning at zero
  int* cell_tmp = nullptr;
  int* cell_xyz = nullptr;  // by Morton rank: x | y<<10 | z<<20
  int* scan_tmp = nullptr;
  // RGL per-atom densities
  double* eb = nullptr;
  // fixed-point positions of the mixed path (null in FP64) and their scale
  uint4* qpos = nullptr;
  double3 qscale;  // 2^32 / L per axis
  // reductions
  double* red_part = nullptr;  // partial sums
  double* red_out = nullptr;   // device scalars
  double* h_red = nullptr;     // pinned host mirror
  int red_blocks = 0;
  DevErr* derr = nullptr;
  DevErr* herr = nullptr;  // pinned
  // energies
  double pe = 0, pe_prev = 0, ke = 0, temp = 0, sim_time = 0;
  std::vector<double> e_nhc;
  bool sorted_once = false;
  bool need_sort = false;  // the next step re-sorts and rebuilds every list (after a restore)
  bool need_forces = false;  // forces are stale: positions were set from the host
  double4* xfer = nullptr;   // [2N] staging in original order (pfm_copy_state, pfm_force_rmse)
  // nvms blocks (pfm_set_speculation): per-step energies and states of a block
  bool speculate = true;
  int spec_k = 0;                        // ring slots
  double* spec_e = nullptr;              // device [spec_k][n_inter + 1]: interaction energies, ke
  double4 *sp_pos = nullptr, *sp_vel = nullptr, *sp_frc = nullptr;  // [spec_k][N]
  uint32_t* sp_mask = nullptr;
  int* sp_orig = nullptr;
  // tuning
  int lanes = 8, block = 256;  // lanes per list row and workgroup size of the pair passes
  // graphene on a metal (rjl on the metal, tb on the graphene, ljc/morsec
  // between them, no common atom): passes over disjoint atoms share launches
  struct GmPlan {
    bool ok = false;
    int rjl = -1, tb = -1, cos = -1;
    bool first_ok = false;  // the rjl rows and the ljc/morsec graphene rows share no atom
  } gm;
  bool fuse = true;  // pfm_set_fusion
  // lj / lj1g / rjl passes grouped into launches (phases) whose passes write
  // disjoint atoms; planned at the first force evaluation
  struct MultiPlan {
    bool planned = false, ok = false;
    // first_ok: every force pass is its atoms' first writer in launch order
    // or writes only atoms an earlier pass wrote (no mixed group)
    bool first_ok = false;
    struct Pass {
      int inter, kind, list;  // kind: 0 lj1g, 1 lj rows of list `list`, 2 rjl density, 3 rjl force
      bool first;             // a force pass whose atoms no earlier pass writes
    };
    std::vector<std::vector<Pass>> phases;
  } mp;
  // captured steady-state blocks of steps (pfm_run)
  bool use_graphs = true;
  int step_base = 0;  // first step of the block being captured (kernel step arguments are relative to it)
  struct GraphEntry {
    std::vector<uint64_t> key;    // configuration and buffer identities at the block start
    std::vector<uint64_t> after;  // buffer identities after the block
    hipGraphExec_t exec;
    long used = 0;  // graph_clock at the last launch
  };
  std::vector<GraphEntry> graphs;
  long graph_captures = 0, graph_launches = 0, graph_clock = 0;
  // runtime group changes (pfm_add_group_change)
  struct GroupChange {
    int from, to, ts1, ts2, frec;
  };
  std::vector<GroupChange> changes;
  // output records (pfm_set_records)
  struct Records {
    std::vector<int> eper, xper;  // energy and frame periods
    pfm_record_fn fn = nullptr;
    void* user = nullptr;
    int xslots = 0;          // frame slots asked for
    int ecap = 0, xcap = 0;  // device/host slots of each ring
    int ewidth = 0;          // doubles per energy slot
    double* h_e = nullptr;                  // pinned [ecap][ewidth]
    double* h_x = nullptr;                  // pinned [xcap][6 n]: positions then velocities
    double* d_x = nullptr;                  // device staging of one frame [6 n]
    double* d_part = nullptr;               // out-block partial sums [11][red_blocks]
    int* d_ctr = nullptr;  // records issued (energy, frame) and the current slots
    long ne = 0, nx = 0;   // host mirror of the issued counts
    struct Pending {
      int step, integ;
      double time;
      int eslot, xslot;  // -1 when absent
      hipEvent_t done;
      std::vector<int> group_n;  // active atoms of every group at the step
    };
    std::deque<Pending> q;
    std::vector<hipEvent_t> free_ev;
    int pending_e = 0, pending_x = 0;
    int fn_rc = 0;
  } rec;
  // timing
  bool timing = false;
  std::vector<KTimer> timers;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  std::string err;
};

#define HIPCHK(h, x)                                                                       \
  do {                                                                                     \
    hipError_t _e = (x);                                                                   \
    if (_e != hipSuccess) {                                                                \
      (h)->err = std::string("error: HIP ") + hipGetErrorString(_e) + " at " #x;           \
      return PFM_E_HIP;                                                                    \
    }                                                                                      \
  } while (0)

// launch bookkeeping (pfm_context.hip)
void tim_begin(pfm_ctx* h, const char* name, hipEvent_t* a, hipEvent_t* b);
void tim_end(pfm_ctx* h, hipEvent_t a, hipEvent_t b, const char* name);

// Timed launches (pfm_set_timing) go through hipExtLaunchKernelGGL, whose
// start and stop events take the timestamps of the kernel's own dispatch, so
// a mean duration excludes the gaps between launches.
#define PFM_LAUNCH(h, name, kern, grid, blk, ...)                                                       \
  do {                                                                                                  \
    if ((h)->timing) {                                                                                  \
      hipEvent_t _a = nullptr, _b = nullptr;                                                            \
      tim_begin((h), name, &_a, &_b);                                                                   \
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(blk), 0, (h)->stream, _a, _b, 0, __VA_ARGS__);       \
      tim_end((h), _a, _b, name);                                                                       \
    } else {                                                                                            \
      hipLaunchKernelGGL(kern, dim3(grid), dim3(blk), 0, (h)->stream, __VA_ARGS__);                     \
    }                                                                                                   \
  } while (0)

static inline int pad8(int nb) { return ((nb + 7) / 8) * 8; }

// kernels' host launchers
int launch_sort_and_bin(pfm_ctx* h, int step, bool permute);
int launch_build_list(pfm_ctx* h, int ii, int k, int step);
int launch_build_pair(pfm_ctx* h, int ii, int k, int jj, int kk, int step);
int launch_build_quad(pfm_ctx* h, int ii, int k, int jj, int kk, int b1i, int b1k, int b2i, int b2k, int step);
int launch_nn_from_tb(pfm_ctx* h, int ii, int step);
int launch_forces(pfm_ctx* h, bool energy, bool first_writes);
// the shared lj / lj1g / rjl launches can write every atom's force first
// (no zeroing before them): see plan_multi
bool multi_first_writes(pfm_ctx* h);
// the co-scheduled launches of graphene on a metal can write every atom's
// force first: see gm_forces
bool gm_first_writes(pfm_ctx* h);
void launch_group_rows(pfm_ctx* h);
void scan_excl(pfm_ctx* h, int n, const int* in, int* out, int* tmp);
int sur_alloc(pfm_ctx* h, Inter& it);
void sur_free(Inter& it);
void launch_sur_converse(pfm_ctx* h, Inter& it, int step);
int launch_sur_forces(pfm_ctx* h, Inter& it, int accumulate, bool keep = false);
int launch_rebo(pfm_ctx* h, Inter& it, bool energy, int accumulate);
// interactions with a converse list nl(2) built with nl(1)'s cutoff and period
static inline bool has_converse(int kind) {
  return kind == PFM_LJ || kind == PFM_LJC || kind == PFM_MORSEC || kind == PFM_SUR;
}
