// pfm_forces.hip -- force and energy passes on gfx950.
//
// Every pass is a gather over full neighbour lists: T lanes share one atom
// row, each lane walks every T-th entry, the row's force is summed across the
// T lanes with DPP/shuffle reductions and written by the row's first lane.
// No atomics; per-atom sums are FP64 in both precisions.  Template parameter
// R is the pair arithmetic (double, or float for the mixed path: dr formed in
// FP64 and rounded once), E adds the per-row energy of the reference's
// *_energy routines.
//
//   rjl     density pass + embedding-force pass   (RosatoGuillopeLegrand.f90:23-94)
//   lj      pair gather over nl(1) and nl(2)      (LennardJones.f90:23-69)
//   lj1g    pair gather over the full list with the reference's half-list
//           weights                                (LennardJones_1g.f90:28-117)
//   tb      bond-order pass + force pass          (TersoffBrenner.f90:26-150)
//   ljc/morsec  normals, graphene pass, normal-derivative pass, metal pass
//           (graphenenorm.f90:38-56, LennardJonesCosine.f90, MorseCosine.f90)
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "pfm_internal.h"

struct RowArgs {
  int nrows, stride;
  const int* rows;
  const int* nnum;
  const int* nbr;
  const double4* pos;  // FP64 positions
  const uint4* qpos;   // fixed-point positions (mixed path)
  float3 hq;           // L / 2^32 per axis
  float hiso;          // L / 2^32 when the box is a cube, else 0 (one replica; else RepDev::hiso)
  double3 hb, L, iL;
  const int* rep;      // replica of every internal atom (null: one replica, geometry above)
  const RepDev* rdev;
  double4* frc;
  int accumulate;  // 0: this pass writes the first contribution of the step
  double* e_row;
};

// geometry of the replica that holds atom i
template <class R> __device__ __forceinline__ Geo<R> geo_of(const RowArgs& a, int i);
template <> __device__ __forceinline__ Geo<double> geo_of<double>(const RowArgs& a, int i) {
  if (a.rep) {
    const RepDev& R = a.rdev[a.rep[i]];
    return Geo<double>{a.pos, R.L, R.iL};
  }
  return Geo<double>{a.pos, a.L, a.iL};
}
template <> __device__ __forceinline__ Geo<float> geo_of<float>(const RowArgs& a, int i) {
  if (a.rep) {
    const float3 hq = a.rdev[a.rep[i]].hq;
    return Geo<float>{a.qpos, hq.x, hq.y, hq.z};
  }
  return Geo<float>{a.qpos, a.hq.x, a.hq.y, a.hq.z};
}

// store v in the w slot of atom i's position record (pos4 or qpos)
template <class R> __device__ __forceinline__ void store_w(const RowArgs& a, int i, double v);
template <> __device__ __forceinline__ void store_w<double>(const RowArgs& a, int i, double v) {
  const_cast<double4*>(a.pos)[i].w = v;
}
template <> __device__ __forceinline__ void store_w<float>(const RowArgs& a, int i, double v) {
  const_cast<uint4*>(a.qpos)[i].w = __float_as_uint((float)v);
}

#define ROW_PROLOGUE(T) ROW_PROLOGUE_AT(T, xcd_block(nb_real))
#define ROW_PROLOGUE_AT(T, LB)                            \
  int lb = (LB);                                          \
  if (lb < 0) return;                                     \
  int gt = lb * blockDim.x + threadIdx.x;                 \
  int row = gt / (T), t = threadIdx.x & ((T)-1);          \
  if (row >= a.nrows) return;                             \
  int i = a.rows[row];                                    \
  const Geo<R> G = geo_of<R>(a, i);                       \
  const typename Geo<R>::P pi = G.at(i);                  \
  int n = a.nnum[row];                                    \
  const int* nbrow = a.nbr + (size_t)row * a.stride;

// Walk a row's entries four at a time: lane t of the row's T lanes takes
// entries 4t..4t+3 of every block of 4T, with one 16-byte index load (rows
// start on 16 bytes and are padded past their end), then the four
// position-record gathers, then the arithmetic f(j, record, valid).  A block
// waits on two memory latencies.  Entries past the end of the row gather the
// row atom's own record and reach f with valid = false (f masks their
// contribution without branching).
// The FP64 passes are bound by their arithmetic; they keep one entry per
// iteration (every T-th entry), which costs fewer registers.
// The FP32 walk calls flush() after every block of four entries: passes that
// sum a block's contributions in FP32 registers (RowAcc<float>) move them to
// the row's FP64 sums there, one conversion and FP64 add per block and
// component instead of one per entry.
struct NoFlush {
  __device__ __forceinline__ void operator()() const {}
};
template <int T, class R, class F, class B = NoFlush>
__device__ __forceinline__ void walk_row(const Geo<R>& G, const int* __restrict__ nbrow, int n, int t, int self, F f,
                                         B flush = B{}) {
  if constexpr (sizeof(R) == 8) {
    for (int p = t; p < n; p += T) {
      const int j = nbrow[p];
      f(j, G.at(j), true);
    }
    return;
  }
  for (int p0 = 4 * t; p0 < n; p0 += 4 * T) {
    const int4 v = *reinterpret_cast<const int4*>(nbrow + p0);
    const int jj[4] = {v.x, p0 + 1 < n ? v.y : self, p0 + 2 < n ? v.z : self, p0 + 3 < n ? v.w : self};
    typename Geo<R>::P q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = G.at(jj[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) f(jj[u], q[u], p0 + u < n);
    flush();
  }
}

// Per-lane sums of a row pass, a vector (fx, fy, fz) and a scalar e:
// force(c, dr) subtracts c dr, add(v) / sub(v) add / subtract the vector v,
// energy(x) adds x to e.  FP64: every term goes straight into the FP64 sums.
// FP32 terms: a block of four entries is summed in FP32 and flush() adds the
// block to the FP64 sums.  The two sides of a pair then round in different
// blocks, so a system's forces sum to zero within about 1e-8 of sum |F_i|
// instead of within FP64 rounding.
template <class R> struct RowAcc;
template <> struct RowAcc<double> {
  double fx = 0, fy = 0, fz = 0, e = 0;
  __device__ __forceinline__ void force(double c, double dx, double dy, double dz) {
    fx -= c * dx; fy -= c * dy; fz -= c * dz;
  }
  __device__ __forceinline__ void add(double x, double y, double z) { fx += x; fy += y; fz += z; }
  __device__ __forceinline__ void sub(double x, double y, double z) { fx -= x; fy -= y; fz -= z; }
  __device__ __forceinline__ void energy(double v) { e += v; }
  __device__ __forceinline__ void flush() {}
};
template <> struct RowAcc<float> {
  double fx = 0, fy = 0, fz = 0, e = 0;
  float bx = 0, by = 0, bz = 0, be = 0;
  __device__ __forceinline__ void force(float c, float dx, float dy, float dz) {
    bx = __builtin_fmaf(-c, dx, bx); by = __builtin_fmaf(-c, dy, by); bz = __builtin_fmaf(-c, dz, bz);
  }
  __device__ __forceinline__ void add(float x, float y, float z) { bx += x; by += y; bz += z; }
  __device__ __forceinline__ void sub(float x, float y, float z) { bx -= x; by -= y; bz -= z; }
  __device__ __forceinline__ void energy(float v) { be += v; }
  __device__ __forceinline__ void flush() {
    fx += (double)bx; fy += (double)by; fz += (double)bz; e += (double)be;
    bx = by = bz = be = 0.0f;
  }
};

// walk_row with one per-atom side value of each neighbour (aux[j]) gathered
// together with its position record: f(j, record, aux value, valid)
template <int T, class R, class A, class F, class B = NoFlush>
__device__ __forceinline__ void walk_row_aux(const Geo<R>& G, const int* __restrict__ nbrow, int n, int t, int self,
                                             const A* __restrict__ aux, F f, B flush = B{}) {
  if constexpr (sizeof(R) == 8) {
    for (int p = t; p < n; p += T) {
      const int j = nbrow[p];
      f(j, G.at(j), aux[j], true);
    }
    return;
  }
  for (int p0 = 4 * t; p0 < n; p0 += 4 * T) {
    const int4 v = *reinterpret_cast<const int4*>(nbrow + p0);
    const int jj[4] = {v.x, p0 + 1 < n ? v.y : self, p0 + 2 < n ? v.z : self, p0 + 3 < n ? v.w : self};
    typename Geo<R>::P q[4];
    A x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      q[u] = G.at(jj[u]);
      x[u] = aux[jj[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) f(jj[u], q[u], x[u], p0 + u < n);
    flush();
  }
}

// walk_row over entries that carry a two-bit tag in bits 30-31 (lj1g's
// weights): f(record, tag, valid)
#define ENTRY_ATOM 0x3fffffff
template <int T, class R, class F, class B = NoFlush>
__device__ __forceinline__ void walk_row_tag(const Geo<R>& G, const int* __restrict__ nbrow, int n, int t, int self,
                                             F f, B flush = B{}) {
  if constexpr (sizeof(R) == 8) {
    for (int p = t; p < n; p += T) {
      const unsigned e = (unsigned)nbrow[p];
      f(G.at((int)(e & ENTRY_ATOM)), e >> 30, true);
    }
    return;
  }
  for (int p0 = 4 * t; p0 < n; p0 += 4 * T) {
    const int4 v = *reinterpret_cast<const int4*>(nbrow + p0);
    const unsigned ee[4] = {(unsigned)v.x, p0 + 1 < n ? (unsigned)v.y : (unsigned)self,
                            p0 + 2 < n ? (unsigned)v.z : (unsigned)self, p0 + 3 < n ? (unsigned)v.w : (unsigned)self};
    typename Geo<R>::P q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = G.at((int)(ee[u] & ENTRY_ATOM));
#pragma unroll
    for (int u = 0; u < 4; ++u) f(q[u], ee[u] >> 30, p0 + u < n);
    flush();
  }
}

__device__ __forceinline__ void write_force(const RowArgs& a, int i, double fx, double fy, double fz) {
  double4 f = a.accumulate ? a.frc[i] : make_double4(0, 0, 0, 0);
  f.x += fx; f.y += fy; f.z += fz;
  a.frc[i] = f;
}

// ------------------------------------------------------------------- RGL

// Mixed-path dr of the RGL passes and its square.  Any box: the fixed-point
// differences scaled per axis (Geo<float>::d2).  ISO (the row's replica is a
// cube, h = L / 2^32 on every axis): the unscaled differences and
// h^2 (ux^2 + uy^2 + uz^2), two multiplies fewer; dr = h (ux, uy, uz), so
// a pass that sums c dr sums c u and multiplies the row's sum by h.  The
// choice is the replica's own, so a replica takes the same arithmetic in a
// batch as alone.
__device__ __forceinline__ float hiso_of(const RowArgs& a, int i) { return a.rep ? a.rdev[a.rep[i]].hiso : a.hiso; }
template <bool ISO>
__device__ __forceinline__ float rjl_d2(const Geo<float>& G, const uint4& a, const uint4& b, float h2, float& dx,
                                        float& dy, float& dz) {
  if constexpr (ISO) {
    dx = (float)(int)(b.x - a.x);
    dy = (float)(int)(b.y - a.y);
    dz = (float)(int)(b.z - a.z);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) * h2;
  } else {
    return G.d2(a, b, dx, dy, dz);
  }
}

struct RjlP {
  double A0, xi, p, q, r0, R1, R2, invw;
};

// largest workgroup of the row passes: FP32 pair arithmetic fits 128 VGPRs
// (16 waves per CU), FP64 needs up to 256
template <class R> struct PairBlock { static constexpr int max = sizeof(R) == 4 ? 1024 : 256; };
#define PAIR_BLOCK (PairBlock<R>::max)

// density pass: rho_i = sum_{r <= R2} exp(-2q(r/r0-1)) fc(r); stores
// Eb_i = sqrt(rho_i) and 1/Eb_i in the w slot of atom i's position record
template <int T, class R>
__device__ __forceinline__ void rjl_density_rows(const RowArgs& a, const RjlP& P, double* __restrict__ eb, int blk) {
  ROW_PROLOGUE_AT(T, blk)
  double rho = 0.0;
  if constexpr (sizeof(R) == 4) {
    // exp(-2q(r/r0 - 1)) = 2^(r kr + k0); cutoff argument in revolutions as in cutoff_bl<float>
    const float R2 = P.R2, ca = 0.5 * P.invw, cb = -0.5 * P.R1 * P.invw;
    const float kr = PM<float>::kexp(-2.0 * P.q / P.r0), k0 = PM<float>::kexp(2.0 * P.q);
    const float hiso = hiso_of(a, i);
    auto body = [&](auto iso) {
      constexpr bool ISO = decltype(iso)::value;
      const float h2 = hiso * hiso;
      RowAcc<float> acc;
      walk_row<T, float>(G, nbrow, n, t, i, [&](int, const uint4& q, bool valid) {
        float dx, dy, dz;
        const float r = __builtin_amdgcn_sqrtf(rjl_d2<ISO>(G, pi, q, h2, dx, dy, dz));
        const float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fmed3f(__builtin_fmaf(r, ca, cb), 0.0f, 0.5f));
        const float e = __builtin_amdgcn_exp2f(__builtin_fmaf(r, kr, k0)) * __builtin_fmaf(0.5f, c, 0.5f);
        acc.energy(valid && r <= R2 ? e : 0.0f);
      }, [&] { acc.flush(); });
      return acc.e;
    };
    rho = hiso > 0.0f ? body(std::true_type{}) : body(std::false_type{});
  } else {
    const R R1 = P.R1, R2 = P.R2, invw = P.invw, ir0 = 1.0 / P.r0, kdf = -0.5 * PFM_PI * P.invw;
    const R kq = PM<R>::kexp(-2.0 * P.q);
    walk_row<T, R>(G, nbrow, n, t, i, [&](int, const typename Geo<R>::P& q, bool valid) {
      R dx, dy, dz;
      R r = PM<R>::sqrt_(G.d2(pi, q, dx, dy, dz));
      R fc, dfr;
      cutoff_bl<R>(r, R1, invw, kdf, fc, dfr);
      R e = PM<R>::expk(kq * (r * ir0 - R(1))) * fc;
      rho += (double)(valid && r <= R2 ? e : R(0));
    });
  }
  rho = group_sum<T>(rho);
  if (t == 0) {
    double s = sqrt(rho);
    eb[i] = s;
    store_w<R>(a, i, 1.0 / s);
  }
}

template <int T, class R>
__global__ void __launch_bounds__(PAIR_BLOCK) k_rjl_density(RowArgs a, RjlP P, double* __restrict__ eb, int nb_real) {
  rjl_density_rows<T, R>(a, P, eb, xcd_block(nb_real));
}

// embedding-force pass (RosatoGuillopeLegrand.f90:80-93):
// F_i -= sum_j [2 A0 (p/r0 fc - df r) e_p - xi (q/r0 fc - df r/2)(1/Eb_i + 1/Eb_j) e_q]/r dr
// with 1/Eb read from the w slot of the gathered position records
template <int T, class R, bool E>
__device__ __forceinline__ void rjl_force_rows(const RowArgs& a, const RjlP& P, const double* __restrict__ eb,
                                               int blk) {
  ROW_PROLOGUE_AT(T, blk)
  double fx = 0, fy = 0, fz = 0, er = 0;
  if constexpr (sizeof(R) == 4) {
    // The two brackets with fc = (1 + cos)/2 and df r = kdf sin folded into
    // their constants: 2 A0 (p/r0 fc - df r) = Ac (1 + cos) + As sin and
    // xi (q/r0 fc - df r/2) = Bc (1 + cos) + Bs sin; the exponentials as
    // 2^(r k + k0).
    const double kdf = -0.5 * PFM_PI * P.invw;
    const float R2 = P.R2, ca = 0.5 * P.invw, cb = -0.5 * P.R1 * P.invw;
    const float Ac = P.A0 * P.p / P.r0, As = -2.0 * P.A0 * kdf;
    const float Bc = 0.5 * P.xi * P.q / P.r0, Bs = -0.5 * P.xi * kdf;
    const float kpr = PM<float>::kexp(-P.p / P.r0), kp0 = PM<float>::kexp(P.p);
    const float kqr = PM<float>::kexp(-2.0 * P.q / P.r0), kq0 = PM<float>::kexp(2.0 * P.q);
    const float ie = Geo<float>::w(pi);
    const float hiso = hiso_of(a, i);
    auto body = [&](auto iso) {
      constexpr bool ISO = decltype(iso)::value;
      const float h2 = hiso * hiso;
      RowAcc<float> acc;
      walk_row<T, float>(G, nbrow, n, t, i, [&](int, const uint4& q, bool valid) {
        float dx, dy, dz, r, ir;
        dist_and_inverse<float>(rjl_d2<ISO>(G, pi, q, h2, dx, dy, dz), r, ir);
        const float x = __builtin_amdgcn_fmed3f(__builtin_fmaf(r, ca, cb), 0.0f, 0.5f);
        const float sn = __builtin_amdgcn_sinf(x), cs = __builtin_amdgcn_cosf(x);
        const float ep = __builtin_amdgcn_exp2f(__builtin_fmaf(r, kpr, kp0));
        const float eq = __builtin_amdgcn_exp2f(__builtin_fmaf(r, kqr, kq0));
        const float A = __builtin_fmaf(As, sn, __builtin_fmaf(Ac, cs, Ac));
        const float B = __builtin_fmaf(Bs, sn, __builtin_fmaf(Bc, cs, Bc));
        const bool in = valid && r <= R2;
        float c = __builtin_fmaf(A, ep, -(B * (ie + Geo<float>::w(q)) * eq)) * ir;
        c = in ? c : 0.0f;
        acc.force(c, dx, dy, dz);
        if (E) {
          const float e = ep * __builtin_fmaf(0.5f, cs, 0.5f);
          acc.energy(in ? e : 0.0f);
        }
      }, [&] { acc.flush(); });
      // with ISO the sums are of c u: dr = h u
      const double hs = ISO ? (double)hiso : 1.0;
      fx = acc.fx * hs; fy = acc.fy * hs; fz = acc.fz * hs; er = acc.e;
    };
    if (hiso > 0.0f) body(std::true_type{});
    else body(std::false_type{});
  } else {
    const R R1 = P.R1, R2 = P.R2, invw = P.invw, ir0 = 1.0 / P.r0, kdf = -0.5 * PFM_PI * P.invw;
    const R kp = PM<R>::kexp(-P.p), kq = PM<R>::kexp(-2.0 * P.q);
    const R c2A = 2.0 * P.A0, pr0 = P.p / P.r0, qr0 = P.q / P.r0, xi = P.xi;
    const R ie = Geo<R>::w(pi);
    walk_row<T, R>(G, nbrow, n, t, i, [&](int, const typename Geo<R>::P& q, bool valid) {
      R dx, dy, dz;
      R r, ir;
      dist_and_inverse<R>(G.d2(pi, q, dx, dy, dz), r, ir);
      R fc, dfr;
      cutoff_bl<R>(r, R1, invw, kdf, fc, dfr);
      R x = r * ir0 - R(1);
      R ep = PM<R>::expk(kp * x), eq = PM<R>::expk(kq * x);
      R c = (c2A * (pr0 * fc - dfr) * ep - xi * (qr0 * fc - R(0.5) * dfr) * (ie + Geo<R>::w(q)) * eq) * ir;
      bool in = valid && r <= R2;
      c = in ? c : R(0);
      fx -= (double)(c * dx); fy -= (double)(c * dy); fz -= (double)(c * dz);
      if (E) er += in ? (double)(ep * fc) : 0.0;
    });
  }
  fx = group_sum<T>(fx); fy = group_sum<T>(fy); fz = group_sum<T>(fz);
  if (E) er = group_sum<T>(er);
  if (t == 0) {
    write_force(a, i, fx, fy, fz);
    if (E) a.e_row[row] = P.A0 * er - P.xi * eb[i];
  }
}

template <int T, class R, bool E>
__global__ void __launch_bounds__(PAIR_BLOCK) k_rjl_force(RowArgs a, RjlP P, const double* __restrict__ eb,
                                                          int nb_real) {
  rjl_force_rows<T, R, E>(a, P, eb, xcd_block(nb_real));
}

// -------------------------------------------------------------------- LJ

struct LjP {
  double eps4, sig, R1, R2, invw;
};

// (E with a.e_row null: the converse rows nl(2), which carry no energy)
template <int T, class R, bool E>
__device__ __forceinline__ void lj_rows(const RowArgs& a, const LjP& P, int blk) {
  ROW_PROLOGUE_AT(T, blk)
  const R R1 = P.R1, R2 = P.R2, invw = P.invw, sig = P.sig, e4 = P.eps4, kdf = -0.5 * PFM_PI * P.invw;
  RowAcc<R> acc;
  walk_row<T, R>(G, nbrow, n, t, i, [&](int, const typename Geo<R>::P& q, bool valid) {
    R dx, dy, dz;
    R r, ir;
    dist_and_inverse<R>(G.d2(pi, q, dx, dy, dz), r, ir);
    R fc, dfr;
    cutoff_bl<R>(r, R1, invw, kdf, fc, dfr);
    R s = sig * ir, s2 = s * s, V = s2 * s2 * s2;
    R c = e4 * (V * (R(12) * V - R(6)) * fc * ir - V * (V - R(1)) * dfr) * ir;
    bool in = valid && r < R2;
    c = in ? c : R(0);
    acc.force(c, dx, dy, dz);
    if (E) {
      const R v = e4 * V * (V - R(1)) * fc;
      acc.energy(in ? v : R(0));
    }
  }, [&] { acc.flush(); });
  const double fx = group_sum<T>(acc.fx), fy = group_sum<T>(acc.fy), fz = group_sum<T>(acc.fz);
  const double e = E ? group_sum<T>(acc.e) : 0.0;
  if (t == 0) {
    write_force(a, i, fx, fy, fz);
    if (E && a.e_row) a.e_row[row] = e;
  }
}

template <int T, class R, bool E>
__global__ void __launch_bounds__(PAIR_BLOCK) k_lj(RowArgs a, LjP P, int nb_real) {
  lj_rows<T, R, E>(a, P, xcd_block(nb_real));
}

// lj1g: the reference sums entries p <= lessnnum(i) of each row once and
// applies Newton's third law (LennardJones_1g.f90:63-99).  Entry k of row i is
// in that half iff loc(k) < cut(i); the full-list gather weights pair (i,k)
// by [loc(k) < cut(i)] + [loc(i) < cut(k)], which reproduces the reference
// for any group order (for ascending groups the weight is always 1).
struct Lj1gP {
  double c6, c12, c6t6, c12t12, R1, R2, invw;
};

// The weights come from the entries' tags (k_lj1g_tag): bit 30 the pair is in
// i's half list, bit 31 in j's.
template <int T, class R, bool E>
__device__ __forceinline__ void lj1g_rows(const RowArgs& a, const Lj1gP& P, int blk) {
  ROW_PROLOGUE_AT(T, blk)
  const R R1 = P.R1, R2 = P.R2, invw = P.invw, c6 = P.c6, c12 = P.c12, c6t6 = P.c6t6, c12t12 = P.c12t12;
  RowAcc<R> acc;
  walk_row_tag<T, R>(G, nbrow, n, t, i, [&](const typename Geo<R>::P& q, unsigned tag, bool valid) {
    int wh = valid && (tag & 1u) ? 1 : 0;
    int w = wh + (valid && (tag & 2u) ? 1 : 0);
    R dx, dy, dz;
    R r, invr2;
    R tr;
    if constexpr (sizeof(R) == 4) {
      // r and 1/r^2 from one v_rsq_f32; the clamp to [R1, R2] in one v_med3
      R ir;
      dist_and_inverse<R>(G.d2(pi, q, dx, dy, dz), r, ir);
      invr2 = ir * ir;
      tr = __builtin_amdgcn_fmed3f(r, R1, R2);
    } else {
      r = PM<R>::sqrt_(G.d2(pi, q, dx, dy, dz));
      invr2 = PM<R>::rcp(r * r);
      tr = r < R1 ? R1 : (r > R2 ? R2 : r);
    }
    // f_dfr_cut (cut_off_poly.f90:30-43), dfr without the 1/(R2-R1) factor
    R x = (tr - R1) * invw, x2 = x * x;
    R f = R(1) + x2 * x * (R(-10) + R(15) * x - R(6) * x2);
    R dfr = tr * x2 * (R(-30) + R(60) * x - R(30) * x2);
    R U = invr2 * invr2 * invr2;
    R Fs = U * invr2 * ((c12t12 * U - c6t6) * f - (c12 * U - c6) * dfr);
    R cw = w ? Fs * R(w) : R(0);  // past the row's end r = 0 and Fs is not finite
    acc.force(cw, dx, dy, dz);
    if (E) {
      // cut_off_poly f_cut (cut_off_poly.f90:6-17)
      R fe = (r > R1 && r < R2) ? f : (r <= R1 ? R(1) : R(0));
      const R v = U * (c12 * U - c6) * fe;
      acc.energy(wh ? v : R(0));
    }
  }, [&] { acc.flush(); });
  const double fx = group_sum<T>(acc.fx), fy = group_sum<T>(acc.fy), fz = group_sum<T>(acc.fz);
  const double e = E ? group_sum<T>(acc.e) : 0.0;
  if (t == 0) {
    write_force(a, i, fx, fy, fz);
    if (E) a.e_row[row] = e;
  }
}

template <int T, class R, bool E>
__global__ void __launch_bounds__(PAIR_BLOCK) k_lj1g(RowArgs a, Lj1gP P, int nb_real) {
  lj1g_rows<T, R, E>(a, P, xcd_block(nb_real));
}

// -------------------------------------------------------------------- TB

struct TbP {
  double D, S, beta, r0, delt, a0, c02, d02, R1, R2, invw;
};

// Tersoff-Brenner rows are short (three bonds in graphene): TB_T lanes share
// a row and lane t owns bonds p = t, t + TB_T, ...; the angular sums of a bond
// over the row's other bonds read their geometry from the owning lane
// (shuffles inside the row's lane group).
#define TB_T 4
#define TB_U (PFM_TB_MAXN / TB_T)

// geometry of this lane's bonds: dr, r, cutoff and its derivative/r
template <class R>
__device__ __forceinline__ void tb_own_bonds(const Geo<R>& G, const typename Geo<R>::P& pi, const int* nbrow, int n,
                                             int t, const TbP& P, R* xs, R* ys, R* zs, R* rs, R* fcs, R* dfcs) {
  const R R1 = P.R1, R2 = P.R2, invw = P.invw;
#pragma unroll
  for (int u = 0; u < TB_U; ++u) {
    const int p = t + TB_T * u;
    xs[u] = ys[u] = zs[u] = R(0);
    rs[u] = R(1);
    fcs[u] = dfcs[u] = R(0);
    if (p < n) {
      rs[u] = PM<R>::sqrt_(G.d2(pi, G.at(nbrow[p]), xs[u], ys[u], zs[u]));
      cos_cutoff<R>(rs[u], R1, R2, invw, fcs[u], dfcs[u]);
    }
  }
}

// bond order B(p,i) (TersoffBrenner.f90:81-91) and, per entry, the factors
// of the force pass: Bp = B**(1/delt+1) and
// BT = delt/2 * Bp * fc * D/(S-1) * S * exp(-sqrt(2S) beta (r-r0)/S)
template <class R>
__device__ __forceinline__ void tb_bond_rows(const RowArgs& a, const TbP& P, double* __restrict__ B,
                                             double* __restrict__ Bp, double* __restrict__ BT, int blk) {
  constexpr int T = TB_T;
  ROW_PROLOGUE_AT(T, blk)
  const R R2 = P.R2, a0 = P.a0, c02 = P.c02, d02 = P.d02;
  R xs[TB_U], ys[TB_U], zs[TB_U], rs[TB_U], fcs[TB_U], dfcs[TB_U];
  tb_own_bonds<R>(G, pi, nbrow, n, t, P, xs, ys, zs, rs, fcs, dfcs);
  R b[TB_U];
#pragma unroll
  for (int u = 0; u < TB_U; ++u) b[u] = R(0);
#pragma unroll
  for (int uq = 0; uq < TB_U; ++uq) {
    for (int tq = 0; tq < T; ++tq) {
      const int q = tq + T * uq;
      if (q >= n) break;  // uniform over the row's lanes
      const R xq = __shfl(xs[uq], tq, T), yq = __shfl(ys[uq], tq, T), zq = __shfl(zs[uq], tq, T);
      const R rq = __shfl(rs[uq], tq, T), fq = __shfl(fcs[uq], tq, T);
#pragma unroll
      for (int u = 0; u < TB_U; ++u) {
        const int p = t + T * u;
        if (p < n && q != p && rs[u] < R2 && rq < R2) {
          R cs = R(1) + (xs[u] * xq + ys[u] * yq + zs[u] * zq) * PM<R>::rcp(rs[u] * rq);
          b[u] += fq * (R(1) + c02 / d02 - c02 * PM<R>::rcp(d02 + cs * cs));
        }
      }
    }
  }
  const R ex = R(1.0 / P.delt + 1.0), md = R(-P.delt);
  const R tailc = R(P.D / (P.S - 1.0) * P.S), tailk = R(-sqrt(2.0 * P.S) * P.beta / P.S), r0 = P.r0;
#pragma unroll
  for (int u = 0; u < TB_U; ++u) {
    const int p = t + T * u;
    if (p >= n) continue;
    R bb = rs[u] < R2 ? PM<R>::pow_(R(1) + a0 * b[u], md) : R(0);
    size_t e = (size_t)row * a.stride + p;
    R bp = bb > R(0) ? PM<R>::pow_(bb, ex) : R(0);
    B[e] = (double)bb;
    Bp[e] = (double)bp;
    BT[e] = (double)(R(0.5 * P.delt) * bp * fcs[u] * tailc * PM<R>::exp_(tailk * (rs[u] - r0)));
  }
}

template <class R>
__global__ void __launch_bounds__(256) k_tb_bond(RowArgs a, TbP P, double* __restrict__ B, double* __restrict__ Bp,
                                                 double* __restrict__ BT, int nb_real) {
  tb_bond_rows<R>(a, P, B, Bp, BT, xcd_block(nb_real));
}

// The normal-derivative terms of ljc/morsec for a graphene row (the rows of
// the tb list are the rows of the nearest-neighbour list nl(3)): row i adds,
// for each of its three nearest neighbours j, the vector that j's graphene
// pass stored for i's place in j's nearest-neighbour row (cos_nn_coef).
struct NdArgs {
  const int* nnum;    // nl(3) row lengths
  const int* nbr;     // nl(3) entries
  const int* rev;     // place of the row atom in the neighbour's nl(3) row
  const int* row_of;  // internal atom -> row of the graphene group
  const double4* gC;
  int stride;
};

// force pass (TersoffBrenner.f90:193-242): for each bond (i,j): dB_ij/dr_i
// over i's row, dB_ji/dr_i and i as third body of j's bonds over j's row;
// with ND also the row's normal-derivative terms (lane q < 3 takes the q-th
// nearest neighbour)
template <class R, bool E, bool ND>
__device__ __forceinline__ void tb_force_rows(const RowArgs& a, const TbP& P, const int* __restrict__ row_of,
                                              const int* __restrict__ rev, const double* __restrict__ B,
                                              const double* __restrict__ Bp, const double* __restrict__ BT,
                                              const NdArgs& nd, int blk) {
  constexpr int T = TB_T;
  ROW_PROLOGUE_AT(T, blk)
  const R R1 = P.R1, R2 = P.R2, invw = P.invw, a0 = P.a0, c02 = P.c02, d02 = P.d02;
  const R D_S1 = R(P.D / (P.S - 1.0)), S = P.S, r0 = P.r0;
  const R s2S = R(sqrt(2.0 * P.S) * P.beta), s2oS = R(sqrt(2.0 / P.S) * P.beta), hdel = R(-P.delt * 0.5);
  R xs[TB_U], ys[TB_U], zs[TB_U], rs[TB_U], fcs[TB_U], dfcs[TB_U];
  tb_own_bonds<R>(G, pi, nbrow, n, t, P, xs, ys, zs, rs, fcs, dfcs);
  // dB(p,i)/dr_i over i's other bonds q
  R bx[TB_U], by[TB_U], bz[TB_U];
#pragma unroll
  for (int u = 0; u < TB_U; ++u) bx[u] = by[u] = bz[u] = R(0);
#pragma unroll
  for (int uq = 0; uq < TB_U; ++uq) {
    for (int tq = 0; tq < T; ++tq) {
      const int q = tq + T * uq;
      if (q >= n) break;  // uniform over the row's lanes
      const R xq = __shfl(xs[uq], tq, T), yq = __shfl(ys[uq], tq, T), zq = __shfl(zs[uq], tq, T);
      const R rq = __shfl(rs[uq], tq, T), fq = __shfl(fcs[uq], tq, T), dfq = __shfl(dfcs[uq], tq, T);
#pragma unroll
      for (int u = 0; u < TB_U; ++u) {
        const int p = t + T * u;
        if (p < n && q != p) {
          const R xp = xs[u], yp = ys[u], zp = zs[u], rp = rs[u];
          const R irp2 = PM<R>::rcp(rp * rp);
          R rr = PM<R>::rcp(rp * rq);
          R cosi = (xp * xq + yp * yq + zp * zq) * rr;
          R tt = d02 + (R(1) + cosi) * (R(1) + cosi);
          R c1 = fq * R(2) * a0 * c02 * (R(1) + cosi) * PM<R>::rcp(tt * tt);
          R c2 = dfq * a0 * (R(1) + c02 / d02 - c02 * PM<R>::rcp(tt));
          R irq2 = PM<R>::rcp(rq * rq);
          bx[u] += c1 * ((xp + xq) * rr - cosi * (xp * irp2 + xq * irq2)) + c2 * xq;
          by[u] += c1 * ((yp + yq) * rr - cosi * (yp * irp2 + yq * irq2)) + c2 * yq;
          bz[u] += c1 * ((zp + zq) * rr - cosi * (zp * irp2 + zq * irq2)) + c2 * zq;
        }
      }
    }
  }
  double fx = 0, fy = 0, fz = 0, e = 0;
#pragma unroll
  for (int u = 0; u < TB_U; ++u) {
    const int p = t + T * u;
    if (p >= n) continue;
    const R xp = xs[u], yp = ys[u], zp = zs[u], rp = rs[u];
    const R irp2 = PM<R>::rcp(rp * rp);
    size_t ep = (size_t)row * a.stride + p;
    R bpp = (R)Bp[ep];
    R gx = bx[u] * bpp, gy = by[u] * bpp, gz = bz[u] * bpp;
    int j = nbrow[p];
    int jr = row_of[j];
    int l = rev[ep];
    if (jr < 0 || l < 0) continue;  // asymmetric list: reported by k_rev
    int nj = a.nnum[jr];
    size_t jbase = (size_t)jr * a.stride;
    const R Bl = (R)B[jbase + l], Bpl = (R)Bp[jbase + l];
    // vector from j to i and its cutoff
    const R xl = -xp, yl = -yp, zl = -zp, rl = rp, irl2 = irp2;
    const R fcl = fcs[u], dfcl = dfcs[u];
    const typename Geo<R>::P pj = G.at(j);
    R tx = 0, ty = 0, tz = 0;
    for (int q = 0; q < nj; ++q) {
      if (q == l) continue;
      R xq, yq, zq;
      const R rq = PM<R>::sqrt_(G.d2(pj, G.at(a.nbr[jbase + q]), xq, yq, zq));
      R fq, dfq;
      cos_cutoff<R>(rq, R1, R2, invw, fq, dfq);
      R rr = PM<R>::rcp(rl * rq);
      R cosi = (xl * xq + yl * yq + zl * zq) * rr;
      R tt = d02 + (R(1) + cosi) * (R(1) + cosi);
      R gq = R(2) * a0 * c02 * (R(1) + cosi) * PM<R>::rcp(tt * tt);
      R vx = -xq * rr + cosi * xl * irl2, vy = -yq * rr + cosi * yl * irl2, vz = -zq * rr + cosi * zl * irl2;
      R c1 = Bpl * fq * gq;
      gx += c1 * vx; gy += c1 * vy; gz += c1 * vz;
      // i as third body of bond (j,k)
      R bt = (R)BT[jbase + q];
      R g2 = dfcl * a0 * (R(1) + c02 / d02 - c02 * PM<R>::rcp(tt));
      R g1 = fcl * gq;
      tx += bt * (g1 * vx + xp * g2);
      ty += bt * (g1 * vy + yp * g2);
      tz += bt * (g1 * vz + zp * g2);
    }
    gx *= hdel; gy *= hdel; gz *= hdel;
    fx += (double)tx; fy += (double)ty; fz += (double)tz;
    if (rp < R2) {
      R fc = fcs[u];
      R dff = dfcs[u] * PM<R>::rcp(fc);
      R av = -s2S * (rp - r0);
      R ea = PM<R>::exp_(av), eas = PM<R>::exp_(av * PM<R>::rcp(S));
      R pre = fc * D_S1;
      R Bm = (R)B[ep] + Bl;
      R k1 = (dff - s2S * PM<R>::rcp(rp)) * ea;
      R k2 = (dff - s2oS * PM<R>::rcp(rp)) * Bm * R(0.5);
      R se = S * eas;
      fx += (double)(pre * (xp * k1 - (gx + xp * k2) * se));
      fy += (double)(pre * (yp * k1 - (gy + yp * k2) * se));
      fz += (double)(pre * (zp * k1 - (gz + zp * k2) * se));
      if (E) e += 0.5 * (double)(pre * (ea - Bm * R(0.5) * se));
    }
  }
  if (ND && t < 3 && nd.nnum[row] == 3) {
    const int j = nd.nbr[(size_t)row * nd.stride + t];
    const int l = nd.rev[(size_t)row * 3 + t];
    if (l >= 0 && l <= 2 && nd.row_of[j] >= 0) {  // reported by k_rev / k_nn_from_tb otherwise
      const double4 c = nd.gC[(size_t)j * 3 + l];
      fx -= c.x; fy -= c.y; fz -= c.z;
    }
  }
  fx = group_sum<T>(fx); fy = group_sum<T>(fy); fz = group_sum<T>(fz);
  if (E) e = group_sum<T>(e);
  if (t == 0) {
    write_force(a, i, fx, fy, fz);
    if (E) a.e_row[row] = e;
  }
}

template <class R, bool E>
__global__ void __launch_bounds__(256) k_tb_force(RowArgs a, TbP P, const int* __restrict__ row_of,
                                                  const int* __restrict__ rev, const double* __restrict__ B,
                                                  const double* __restrict__ Bp, const double* __restrict__ BT,
                                                  int nb_real) {
  tb_force_rows<R, E, false>(a, P, row_of, rev, B, Bp, BT, NdArgs{}, xcd_block(nb_real));
}

// ------------------------------------------------------------ ljc / morsec

struct CosP {
  double pref, sig, a, r0, delt, R1, R2, invw, nn_coef;
  int morse;
};

template <class R>
__device__ __forceinline__ void cos_pair(const CosP& P, R r, R ir, R nd, R nnorm_r, R& A, R& Bc, R& S_coef,
                                         R& epair) {
  const R R1 = P.R1, invw = P.invw, delt = P.delt, kdf = -0.5 * PFM_PI * P.invw;
  R fc, dfr;
  cutoff_bl<R>(r, R1, invw, kdf, fc, dfr);
  R dfc = dfr * ir;
  R V2;
  if (P.morse) V2 = PM<R>::exp_(-R(P.a) * (r - R(P.r0)));
  else { R s = R(P.sig) * ir, s2 = s * s; V2 = s2 * s2 * s2; }
  R V1 = V2 * V2;
  R c = nd < R(0) ? -nd : nd;
  R V3 = PM<R>::pow_(c * nnorm_r, delt);
  R ind = PM<R>::rcp(nd);
  if (P.morse) {
    A = (R(2) * (R(P.a) * V1 - (R(P.a) + delt * ir) * V2 * V3) * ir * fc - (V1 - R(2) * V2 * V3) * dfc);
    Bc = R(2) * delt * V2 * V3 * ind * fc;
    epair = V1 - R(2) * V2 * V3;
  } else {
    A = ((R(12) * V1 - (R(6) + delt) * V2 * V3) * ir * ir * fc - (V1 - V2 * V3) * dfc);
    Bc = delt * V2 * V3 * ind * fc;
    epair = V1 - V2 * V3;
  }
  epair *= fc;
  S_coef = V2 * V3 * fc * ind;
}

// Coefficients of the normal-derivative term (*_forces_for_graphene second
// loop, MorseCosine.f90:81-107), per graphene atom j.  For the l-th of j's
// three nearest neighbours i, the reference's sum over j's metal neighbours p
// of  c V2 V3 fc /(den (n_j.dr_jp)) (dr_jp . W) n_j  is linear in dr_jp, so
// it equals  c/den (S_j . W) n_j  with S_j = sum_p V2 V3 fc/(n_j.dr_jp) dr_jp.
// Lanes l = t, t+T, .. of row j store that vector for i = its l-th nearest
// neighbour, in FP64.  The FP64 path takes dr from the FP64 positions; the
// mixed path from the fixed-point records it already reads (the exact wrapped
// difference, in FP64: within two quantisation steps, 2.7e-8 A at L = 115 A).
template <int T, class R>
__device__ __forceinline__ void cos_nn_coef(const RowArgs& a, const CosP& P, const int* __restrict__ nn, int row,
                                            int i, int t, double sx, double sy, double sz, const double4& g,
                                            double4* __restrict__ gC) {
  if (t >= 3) return;
  double d[3][3];
  if constexpr (sizeof(R) == 8) {
    const Geo<double> G = geo_of<double>(a, i);
    const double4 pj = G.at(i);
    for (int k = 0; k < 3; ++k) G.d2(pj, G.at(nn[(size_t)row * 3 + k]), d[k][0], d[k][1], d[k][2]);
  } else {
    const double3 L = a.rep ? a.rdev[a.rep[i]].L : a.L;
    const double hx = L.x / 4294967296.0, hy = L.y / 4294967296.0, hz = L.z / 4294967296.0;
    const uint4 pj = a.qpos[i];
    for (int k = 0; k < 3; ++k) {
      const uint4 q = a.qpos[nn[(size_t)row * 3 + k]];
      d[k][0] = (double)(int)(q.x - pj.x) * hx;
      d[k][1] = (double)(int)(q.y - pj.y) * hy;
      d[k][2] = (double)(int)(q.z - pj.z) * hz;
    }
  }
  for (int l1 = t; l1 < 3; l1 += T) {
    int l2 = (l1 + 1) % 3, l3 = (l1 + 2) % 3;
    double d12[3], d31[3], d23[3];
    for (int k = 0; k < 3; ++k) {
      d12[k] = d[l2][k] - d[l1][k];
      d31[k] = d[l1][k] - d[l3][k];
      d23[k] = d[l3][k] - d[l2][k];
    }
    double s1212 = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
    double s3131 = d31[0] * d31[0] + d31[1] * d31[1] + d31[2] * d31[2];
    double s1231 = d12[0] * d31[0] + d12[1] * d31[1] + d12[2] * d31[2];
    double s2331 = d23[0] * d31[0] + d23[1] * d31[1] + d23[2] * d31[2];
    double s2312 = d23[0] * d12[0] + d23[1] * d12[1] + d23[2] * d12[2];
    double den = s1212 * s3131 - s1231 * s1231;
    double w = sx * (d12[0] * s2331 - d31[0] * s2312) + sy * (d12[1] * s2331 - d31[1] * s2312) +
               sz * (d12[2] * s2331 - d31[2] * s2312);
    double c = P.nn_coef * w / den;
    gC[(size_t)i * 3 + l1] = make_double4(c * g.x, c * g.y, c * g.z, 0);
  }
}

// graphene side, pair and normal terms (*_forces_for_graphene first loop).
// The row's own normal (graphenenorm.f90:38-56) is formed here from its three
// nearest neighbours (rows of nl(3) are the rows of nl(1)) and stored for the
// metal pass (gnorm, and gnf in FP32 for the mixed path); with nearest
// neighbours, the row's normal-derivative coefficients go to gC.
template <int T, class R, bool E>
__device__ __forceinline__ void cos_graphene_rows(const RowArgs& a, const CosP& P, const int* __restrict__ nn,
                                                  int simplified, double4* __restrict__ gnorm,
                                                  float4* __restrict__ gnf, double4* __restrict__ gC, int blk) {
  ROW_PROLOGUE_AT(T, blk)
  double4 g;
  if (simplified) {
    g = make_double4(0, 0, 1, 0);
  } else {
    R d[3][3];
    for (int k = 0; k < 3; ++k) G.d2(pi, G.at(nn[(size_t)row * 3 + k]), d[k][0], d[k][1], d[k][2]);
    R ax = d[1][0] - d[0][0], ay = d[1][1] - d[0][1], az = d[1][2] - d[0][2];
    R bx = d[0][0] - d[2][0], by = d[0][1] - d[2][1], bz = d[0][2] - d[2][2];
    R gx = ay * bz - az * by, gy = az * bx - ax * bz, gz = ax * by - ay * bx;
    if (gz < R(0)) { gx = -gx; gy = -gy; gz = -gz; }
    R inv = R(1) / PM<R>::sqrt_(gx * gx + gy * gy + gz * gz);
    g = make_double4(gx * inv, gy * inv, gz * inv, 0);
  }
  const R nx = g.x, ny = g.y, nz = g.z, R2 = P.R2, pref = P.pref;
  RowAcc<R> acc, sacc;  // force and energy; S
  walk_row<T, R>(G, nbrow, n, t, i, [&](int, const typename Geo<R>::P& q, bool valid) {
    R dx, dy, dz;
    R r, ir;
    dist_and_inverse<R>(G.d2(pi, q, dx, dy, dz), r, ir);
    R nd = nx * dx + ny * dy + nz * dz;
    R A, Bc, Sc, ep;
    cos_pair<R>(P, r, ir, nd, ir, A, Bc, Sc, ep);
    bool in = valid && r < R2;
    A = in ? A * pref : R(0);
    Bc = in ? Bc * pref : R(0);
    Sc = in ? Sc : R(0);
    acc.sub(A * dx + Bc * nx, A * dy + Bc * ny, A * dz + Bc * nz);
    sacc.add(Sc * dx, Sc * dy, Sc * dz);
    if (E) {
      const R v = ep * pref;
      acc.energy(in ? v : R(0));
    }
  }, [&] { acc.flush(); sacc.flush(); });
  const double fx = group_sum<T>(acc.fx), fy = group_sum<T>(acc.fy), fz = group_sum<T>(acc.fz);
  const double e = E ? group_sum<T>(acc.e) : 0.0;
  if (!simplified) {
    const double sx = group_sum<T>(sacc.fx), sy = group_sum<T>(sacc.fy), sz = group_sum<T>(sacc.fz);
    cos_nn_coef<T, R>(a, P, nn, row, i, t, sx, sy, sz, g, gC);
  }
  if (t == 0) {
    write_force(a, i, fx, fy, fz);
    gnorm[i] = g;
    if (sizeof(R) == 4) gnf[i] = make_float4((float)g.x, (float)g.y, (float)g.z, 0.f);
    if (E) a.e_row[row] = e;
  }
}

template <int T, class R, bool E>
__global__ void __launch_bounds__(PAIR_BLOCK) k_cos_graphene(RowArgs a, CosP P, const int* __restrict__ nn,
                                                             int simplified, double4* __restrict__ gnorm,
                                                             float4* __restrict__ gnf, double4* __restrict__ gC,
                                                             int nb_real) {
  cos_graphene_rows<T, R, E>(a, P, nn, simplified, gnorm, gnf, gC, xcd_block(nb_real));
}

// normal-derivative rows (*_forces_for_graphene second loop): row i adds, for
// each of its three nearest neighbours j, the vector that j's graphene pass
// stored for i's place l in j's nearest-neighbour row (cos_nn_coef)
__device__ __forceinline__ void cos_nn_rows(const RowArgs& a, const int* __restrict__ row_of,
                                            const int* __restrict__ rev, const double4* __restrict__ gC, int block) {
  int lb = block;
  int row = lb * blockDim.x + threadIdx.x;
  if (row >= a.nrows) return;
  const int i = a.rows[row];
  const int n = a.nnum[row];
  const int* nbrow = a.nbr + (size_t)row * a.stride;
  int jj[3], ll[3];
  for (int q = 0; q < 3; ++q) {
    jj[q] = n == 3 ? nbrow[q] : 0;
    ll[q] = n == 3 ? rev[(size_t)row * 3 + q] : -1;
  }
  int jr[3];
  double4 c[3];
  for (int q = 0; q < 3; ++q) {
    jr[q] = row_of[jj[q]];
    const bool ok = ll[q] >= 0 && ll[q] <= 2;  // reported by k_rev / k_nn_from_tb otherwise
    c[q] = ok ? gC[(size_t)jj[q] * 3 + ll[q]] : make_double4(0, 0, 0, 0);
    if (!ok) jr[q] = -1;
  }
  double fx = 0, fy = 0, fz = 0;
  for (int q = 0; q < 3; ++q) {
    if (jr[q] < 0) continue;
    fx -= c[q].x; fy -= c[q].y; fz -= c[q].z;
  }
  write_force(a, i, fx, fy, fz);
}

__global__ void __launch_bounds__(256) k_cos_nn(RowArgs a, const int* __restrict__ row_of, const int* __restrict__ rev,
                                                const double4* __restrict__ gC, int nb_real) {
  int lb = xcd_block(nb_real);
  if (lb < 0) return;
  cos_nn_rows(a, row_of, rev, gC, lb);
}

// metal side (*_forces_for_other_atoms): rows of nl(2), neighbours are
// graphene atoms whose normal enters V3 (morsec divides by |n|); the mixed
// path gathers the FP32 copy of the normals (the same values it rounds to).
// Only the rows with entries run, from the work list k_mark_rows builds with
// the list (most metal atoms lie out of reach of the sheet); the pass always
// adds to forces an earlier pass has written, so the other rows keep theirs.
// Work items take hardware block order, which spreads the short list over
// all XCDs.
template <int T, class R>
__device__ __forceinline__ void cos_metal_rows(const RowArgs& a, const CosP& P, const double4* __restrict__ gnorm,
                                               const float4* __restrict__ gnf, const int* __restrict__ mrows,
                                               int block) {
  const int w = (block * blockDim.x + threadIdx.x) / T, t = threadIdx.x & (T - 1);
  if (w >= mrows[a.nrows]) return;
  const int row = mrows[w];
  const int i = a.rows[row];
  const Geo<R> G = geo_of<R>(a, i);
  const typename Geo<R>::P pi = G.at(i);
  const int n = a.nnum[row];
  const int* nbrow = a.nbr + (size_t)row * a.stride;
  const R R2 = P.R2, pref = P.pref;
  RowAcc<R> acc;
  auto f = [&](int, const typename Geo<R>::P& q, const auto& g, bool valid) {
    R dx, dy, dz;
    R r, ir;
    dist_and_inverse<R>(G.d2(pi, q, dx, dy, dz), r, ir);
    const R gx = g.x, gy = g.y, gz = g.z;
    R nd = gx * dx + gy * dy + gz * dz;
    R nn = P.morse ? PM<R>::sqrt_(gx * gx + gy * gy + gz * gz) * r : r;
    R A, Bc, Sc, ep;
    cos_pair<R>(P, r, ir, nd, P.morse ? PM<R>::rcp(nn) : ir, A, Bc, Sc, ep);
    bool in = valid && r < R2;
    A = in ? A * pref : R(0);
    Bc = in ? Bc * pref : R(0);
    acc.sub(A * dx + Bc * gx, A * dy + Bc * gy, A * dz + Bc * gz);
  };
  auto flush = [&] { acc.flush(); };
  if constexpr (sizeof(R) == 4) walk_row_aux<T, R>(G, nbrow, n, t, i, gnf, f, flush);
  else walk_row_aux<T, R>(G, nbrow, n, t, i, gnorm, f, flush);
  const double fx = group_sum<T>(acc.fx), fy = group_sum<T>(acc.fy), fz = group_sum<T>(acc.fz);
  if (t == 0) write_force(a, i, fx, fy, fz);
}

template <int T, class R>
__global__ void __launch_bounds__(PAIR_BLOCK) k_cos_metal(RowArgs a, CosP P, const double4* __restrict__ gnorm,
                                                          const float4* __restrict__ gnf,
                                                          const int* __restrict__ mrows, int nb_real) {
  if ((int)blockIdx.x >= nb_real) return;
  cos_metal_rows<T, R>(a, P, gnorm, gnf, mrows, blockIdx.x);
}

// the normal-derivative rows (graphene atoms) and the metal rows in one
// launch, for groups without common atoms (every force is written once):
// blocks [0, nb_nn) take the graphene rows, the rest the metal work list
template <int T, class R>
__global__ void __launch_bounds__(256) k_cos_nn_metal(RowArgs an, RowArgs am, CosP P, const int* __restrict__ row_of,
                                                      const int* __restrict__ rev, const double4* __restrict__ gC,
                                                      const double4* __restrict__ gnorm, const float4* __restrict__ gnf,
                                                      const int* __restrict__ mrows, int nb_nn, int nb_real) {
  const int lb = blockIdx.x;
  if (lb >= nb_real) return;
  if (lb < nb_nn) cos_nn_rows(an, row_of, rev, gC, lb);
  else cos_metal_rows<T, R>(am, P, gnorm, gnf, mrows, lb - nb_nn);
}

// ------------------------------------------------ passes sharing a launch
//
// lj, lj1g and rjl passes whose rows write disjoint atoms (the species of a
// gas, the replicas of a batch that hold other materials) run as sub-ranges
// of one launch: a binary gas's lj1g A and B rows and a batch's rjl
// densities together, then the lj rows of both directions and the rjl
// forces.  Block b belongs to the last sub-range starting at or before it;
// each sub-range is XCD-remapped on its own.

#define MULTI_MAXP 6
enum MultiKind { MK_LJ1G = 0, MK_LJ = 1, MK_RJL_DENSITY = 2, MK_RJL_FORCE = 3 };
struct MultiPass {
  int kind, start, np, nb;  // first block of the sub-range, its padded and real block counts
  RowArgs a;
  union {
    RjlP rjl;
    LjP lj;
    Lj1gP lj1g;
  } p;
  double* eb;
};
struct MultiArgs {
  int n;
  MultiPass s[MULTI_MAXP];
};

template <int T, class R, bool E>
__global__ void __launch_bounds__(256) k_multi(MultiArgs A) {
  const int b = blockIdx.x;
  int k = 0;
  while (k + 1 < A.n && b >= A.s[k + 1].start) ++k;
  const MultiPass& S = A.s[k];
  const int lb = xcd_local(b - S.start, S.np, S.nb);
  if (lb < 0) return;
  switch (S.kind) {
    case MK_LJ1G: lj1g_rows<T, R, E>(S.a, S.p.lj1g, lb); break;
    case MK_LJ: lj_rows<T, R, E>(S.a, S.p.lj, lb); break;
    case MK_RJL_DENSITY: rjl_density_rows<T, R>(S.a, S.p.rjl, S.eb, lb); break;
    default: rjl_force_rows<T, R, E>(S.a, S.p.rjl, S.eb, lb); break;
  }
}

// ------------------------------------------------- graphene on a metal
//
// With rjl on the metal, tb on the graphene and ljc/morsec between them (no
// common atom), the passes of the two sides share launches: the graphene
// passes are short rows (three bonds, or about forty metal entries) whose
// blocks wait on chains of dependent loads, and on their own they fill a
// fraction of the GPU.  Three launches replace six:
//   first:  tb bond orders (graphene rows) | rjl densities (metal rows)
//   second: ljc/morsec graphene pass       | rjl forces
//   third:  tb forces + normal-derivative  | ljc/morsec metal rows
// The graphene range comes first in each grid, so its long blocks start
// early; each range is XCD-remapped on its own (the metal work list keeps
// hardware order).  Every atom's force is summed by one row per launch.

template <int T, class R>
__global__ void __launch_bounds__(256) k_gm_first(RowArgs ac, TbP Pt, double* __restrict__ B, double* __restrict__ Bp,
                                                  double* __restrict__ BT, int npc, int nbc, RowArgs am, RjlP Pr,
                                                  double* __restrict__ eb, int npm, int nbm) {
  const int b = blockIdx.x;
  if (b < npc) {
    const int lb = xcd_local(b, npc, nbc);
    if (lb >= 0) tb_bond_rows<R>(ac, Pt, B, Bp, BT, lb);
  } else {
    const int lb = xcd_local(b - npc, npm, nbm);
    if (lb >= 0) rjl_density_rows<T, R>(am, Pr, eb, lb);
  }
}

template <int T, class R, bool E>
__global__ void __launch_bounds__(256) k_gm_second(RowArgs ag, CosP Pc, const int* __restrict__ nn, int simplified,
                                                   double4* __restrict__ gnorm, float4* __restrict__ gnf,
                                                   double4* __restrict__ gC, int npc, int nbc, RowArgs am, RjlP Pr,
                                                   const double* __restrict__ eb, int npm, int nbm) {
  const int b = blockIdx.x;
  if (b < npc) {
    const int lb = xcd_local(b, npc, nbc);
    if (lb >= 0) cos_graphene_rows<(T < 4 ? T : 4), R, E>(ag, Pc, nn, simplified, gnorm, gnf, gC, lb);
  } else {
    const int lb = xcd_local(b - npc, npm, nbm);
    if (lb >= 0) rjl_force_rows<T, R, E>(am, Pr, eb, lb);
  }
}

template <int T, class R, bool E, bool ND>
__global__ void __launch_bounds__(256) k_gm_third(RowArgs at, TbP Pt, const int* __restrict__ row_of,
                                                  const int* __restrict__ rev, const double* __restrict__ B,
                                                  const double* __restrict__ Bp, const double* __restrict__ BT,
                                                  NdArgs nd, int npc, int nbc, RowArgs amet, CosP Pc,
                                                  const double4* __restrict__ gnorm, const float4* __restrict__ gnf,
                                                  const int* __restrict__ mrows, int nbm) {
  const int b = blockIdx.x;
  if (b < npc) {
    const int lb = xcd_local(b, npc, nbc);
    if (lb >= 0) tb_force_rows<R, E, ND>(at, Pt, row_of, rev, B, Bp, BT, nd, lb);
  } else if (b - npc < nbm) {
    cos_metal_rows<T, R>(amet, Pc, gnorm, gnf, mrows, b - npc);
  }
}

// ----------------------------------------------------------------- dispatch

static RowArgs row_args(pfm_ctx* h, Inter& it, int k, int accumulate) {
  ListDev& L = it.nl[k];
  RowArgs a;
  a.nrows = L.nrows;
  a.stride = L.stride;
  a.rows = h->grp.rows[L.g1];
  a.nnum = L.nnum;
  a.nbr = L.nbr;
  a.pos = h->pos4;
  a.qpos = h->qpos;
  a.hq = make_float3((float)(h->box[0] / 4294967296.0), (float)(h->box[1] / 4294967296.0),
                     (float)(h->box[2] / 4294967296.0));
  a.hb = make_double3(h->hbox[0], h->hbox[1], h->hbox[2]);
  a.L = make_double3(h->box[0], h->box[1], h->box[2]);
  a.iL = make_double3(1.0 / h->box[0], 1.0 / h->box[1], 1.0 / h->box[2]);
  a.rep = h->rep;
  a.rdev = h->rdev;
  a.hiso = (a.hq.x == a.hq.y && a.hq.y == a.hq.z) ? a.hq.x : 0.0f;
  a.frc = h->frc4;
  a.accumulate = accumulate;
  a.e_row = it.e_row;
  return a;
}

static int row_blocks(int nrows, int T, int block) { return (int)(((long)nrows * T + block - 1) / block); }

// lanes per row of an interaction's separately launched row passes (the
// co-scheduled launches take the handle's count for all their passes)
static int pass_lanes(const pfm_ctx* h, const Inter& it) { return it.lanes ? it.lanes : h->lanes; }

#define DISPATCH_T(T_, ...)                       \
  switch (T_) {                                   \
    case 1: { constexpr int T = 1; __VA_ARGS__; } break;   \
    case 2: { constexpr int T = 2; __VA_ARGS__; } break;   \
    case 4: { constexpr int T = 4; __VA_ARGS__; } break;   \
    case 16: { constexpr int T = 16; __VA_ARGS__; } break; \
    case 32: { constexpr int T = 32; __VA_ARGS__; } break; \
    default: { constexpr int T = 8; __VA_ARGS__; } break;  \
  }

// the lane counts of the co-scheduled launches
#define DISPATCH_T3(T_, ...)                                 \
  switch (T_) {                                            \
    case 4: { constexpr int T = 4; __VA_ARGS__; } break;   \
    case 16: { constexpr int T = 16; __VA_ARGS__; } break; \
    default: { constexpr int T = 8; __VA_ARGS__; } break;  \
  }

static TbP tb_params(const Inter& it) {
  return TbP{it.p[0], it.p[1], it.p[2], it.p[3], it.p[4], it.p[5], it.p[6] * it.p[6], it.p[7] * it.p[7],
             it.p[8], it.p[9], 1.0 / (it.p[9] - it.p[8])};
}
static RjlP rjl_params(const Inter& it) {
  return RjlP{it.p[0], it.p[1], it.p[2], it.p[3], it.p[4], it.p[5], it.p[6], 1.0 / (it.p[6] - it.p[5])};
}
static CosP cos_params(const Inter& it) {
  if (it.kind == PFM_MORSEC)
    return CosP{it.p[0], 0.0, it.p[2], it.p[1], it.p[3], it.p[4], it.p[5], 1.0 / (it.p[5] - it.p[4]),
                2.0 * it.p[0] * it.p[3], 1};
  return CosP{4.0 * it.p[0], it.p[1], 0.0, 0.0, it.p[2], it.p[3], it.p[4], 1.0 / (it.p[4] - it.p[3]),
              4.0 * it.p[0] * it.p[2], 0};
}

static LjP lj_params(const Inter& it) {
  return LjP{4.0 * it.p[0], it.p[1], it.p[2], it.p[3], 1.0 / (it.p[3] - it.p[2])};
}
static Lj1gP lj1g_params(const Inter& it) {
  double eps = it.p[0], sig = it.p[1], s6 = sig * sig * sig * sig * sig * sig;
  return Lj1gP{4.0 * eps * s6, 4.0 * eps * s6 * s6, 6.0 * 4.0 * eps * s6, 12.0 * 4.0 * eps * s6 * s6,
               it.p[2], it.p[3], 1.0 / (it.p[3] - it.p[2])};
}

// Phases of lj / lj1g / rjl passes: each pass goes into the earliest phase
// after its prerequisite (an rjl force after its density) holding no pass
// that writes any of its atoms (forces, or the density's per-atom value).
// Worth it when fewer launches result; other interactions keep separate
// launches (plan not taken).
static void plan_multi(pfm_ctx* h) {
  auto& M = h->mp;
  M.planned = true;
  M.ok = false;
  M.phases.clear();
  struct P {
    pfm_ctx::MultiPlan::Pass pass;
    int group, after;
  };
  std::vector<P> passes;
  for (size_t ii = 0; ii < h->inter.size(); ++ii) {
    const Inter& it = h->inter[ii];
    const int I = (int)ii;
    if (it.kind == PFM_LJ1G) {
      passes.push_back({{I, MK_LJ1G, 0}, it.nl[0].g1, -1});
    } else if (it.kind == PFM_LJ) {
      for (int k = 0; k < 2; ++k) passes.push_back({{I, MK_LJ, k}, it.nl[k].g1, -1});
    } else if (it.kind == PFM_RJL) {
      passes.push_back({{I, MK_RJL_DENSITY, 0}, it.nl[0].g1, -1});
      passes.push_back({{I, MK_RJL_FORCE, 0}, it.nl[0].g1, (int)passes.size() - 1});
    } else {
      return;
    }
  }
  if (passes.size() > MULTI_MAXP) return;
  // groups sharing an atom
  auto overlap = [&](int g1, int g2) {
    if (g1 == g2) return true;
    std::vector<char> in(h->N, 0);
    for (int x : h->grp.idx[g1]) in[x] = 1;
    for (int x : h->grp.idx[g2])
      if (in[x]) return true;
    return false;
  };
  std::vector<int> phase_of(passes.size(), 0);
  int nphase = 0;
  for (size_t a = 0; a < passes.size(); ++a) {
    int ph = passes[a].after >= 0 ? phase_of[passes[a].after] + 1 : 0;
    for (;; ++ph) {
      bool clash = false;
      for (size_t b = 0; b < a && !clash; ++b) {
        if (phase_of[b] != ph) continue;
        // a density and a force pass of one atom set write different things,
        // but the force pass of another interaction would read the
        // density's per-atom value while it is written: keep them apart too
        clash = overlap(passes[a].group, passes[b].group);
      }
      if (!clash) break;
    }
    phase_of[a] = ph;
    nphase = std::max(nphase, ph + 1);
  }
  if (nphase >= (int)passes.size()) return;
  M.phases.assign(nphase, {});
  for (size_t a = 0; a < passes.size(); ++a) M.phases[phase_of[a]].push_back(passes[a].pass);
  // first writers in launch order (phases in turn; the passes of a phase
  // hold disjoint atoms): a force pass is first when none of its group's
  // atoms is written earlier; a force pass with some atoms written earlier
  // and some not leaves the plan without first writes
  std::vector<char> written(h->N, 0);
  M.first_ok = true;
  for (auto& ph : M.phases) {
    for (auto& ps : ph) {
      ps.first = false;
      if (ps.kind == MK_RJL_DENSITY) continue;
      const Inter& it = h->inter[ps.inter];
      const auto& members = h->grp.idx[it.nl[ps.list].g1];
      size_t before = 0;
      for (int x : members) before += written[x] ? 1 : 0;
      if (before == 0) ps.first = true;
      else if (before != members.size()) M.first_ok = false;
    }
    for (auto& ps : ph)
      if (ps.first)
        for (int x : h->grp.idx[h->inter[ps.inter].nl[ps.list].g1]) written[x] = 1;
  }
  M.ok = true;
}

// The shared launches write every atom's force first when the plan has first
// writers only (plan_multi), no group changes, and the first writers' rows
// cover every atom now: their groups are disjoint, so the row counts add up
// to N exactly then.
bool multi_first_writes(pfm_ctx* h) {
  if (!h->mp.planned) plan_multi(h);
  const int T = h->lanes;
  if (!h->fuse || !h->mp.ok || !h->mp.first_ok || !h->changes.empty() || (T != 4 && T != 8 && T != 16)) return false;
  long rows = 0;
  for (const auto& ph : h->mp.phases)
    for (const auto& ps : ph)
      if (ps.first) rows += h->inter[ps.inter].nl[ps.list].nrows;
  return rows == h->N;
}

// the phases of h->mp, one launch each.  Every pass adds to forces the kick
// cleared, or with `first` (multi_first_writes) the first writer of each atom
// writes its force.  Returns 1 when the lane count has no variant here.
template <class R>
static int multi_forces(pfm_ctx* h, bool energy, bool first) {
  const int T = h->lanes;
  if (T != 4 && T != 8 && T != 16) return 1;
  const int BS = 256;
  for (const auto& ph : h->mp.phases) {
    MultiArgs A;
    memset(&A, 0, sizeof A);
    int start = 0;
      for (const auto& ps : ph) {
      Inter& it = h->inter[ps.inter];
      MultiPass& S = A.s[A.n];
      S.kind = ps.kind;
      S.a = row_args(h, it, ps.list, first && ps.first ? 0 : 1);
      if (S.a.nrows == 0) continue;
      if (ps.kind == MK_LJ && ps.list == 1) S.a.e_row = nullptr;  // nl(2) rows carry no energy
      if (ps.kind == MK_LJ1G) S.p.lj1g = lj1g_params(it);
      else if (ps.kind == MK_LJ) S.p.lj = lj_params(it);
      else S.p.rjl = rjl_params(it);
      S.eb = h->eb;
      S.nb = row_blocks(S.a.nrows, T, BS);
      S.np = pad8(S.nb);
      S.start = start;
      start += S.np;
      ++A.n;
    }
    if (A.n == 0) continue;
    DISPATCH_T3(T, {
      if (energy) PFM_LAUNCH(h, "multi", (k_multi<T, R, true>), start, BS, A);
      else PFM_LAUNCH(h, "multi", (k_multi<T, R, false>), start, BS, A);
    });
  }
  return PFM_OK;
}

// The three co-scheduled launches of graphene on a metal (h->gm); every
// pass adds to forces the step's kick cleared, or with `first`
// (gm_first_writes) the second launch writes them.  Returns 1 when the lane
// count has no fused variant (the caller launches the passes one by one).
template <class R>
static int gm_forces(pfm_ctx* h, bool energy, bool first) {
  const int T = h->lanes;
  if (T != 4 && T != 8 && T != 16) return 1;
  Inter& ir = h->inter[h->gm.rjl];
  Inter& it = h->inter[h->gm.tb];
  Inter& ic = h->inter[h->gm.cos];
  if (ir.nl[0].nrows == 0 || it.nl[0].nrows == 0 || ic.nl[0].nrows == 0) return 1;
  const int BS = 256;
  // with `first` the second launch's passes (graphene rows, rjl forces) write
  // their rows' forces and the third launch adds to them
  const int acc2 = first ? 0 : 1;
  RowArgs am = row_args(h, ir, 0, acc2), at = row_args(h, it, 0, 1), ag = row_args(h, ic, 0, acc2);
  RowArgs amet = row_args(h, ic, 1, 1);
  const RjlP Pr = rjl_params(ir);
  const TbP Pt = tb_params(it);
  const CosP Pc = cos_params(ic);
  const bool morse = ic.kind == PFM_MORSEC;
  const int simp = (int)ic.p[morse ? 6 : 5] ? 1 : 0;
  ListDev& NN = ic.nl[2];
  const bool nd = !simp && NN.nrows > 0;
  double* BT = it.Bp + (size_t)it.nl[0].nrows * it.nl[0].stride;
  const int* row_of_tb = h->grp.row_of[it.nl[0].g1];
  NdArgs ndA{NN.nnum, NN.nbr, NN.rev, h->grp.row_of[NN.g1], ic.gC, NN.stride};
  // blocks of each range
  const int nbt = row_blocks(at.nrows, TB_T, BS), npt = pad8(nbt);
  const int nbm = row_blocks(am.nrows, T, BS), npm = pad8(nbm);
  const int Tg = T < 4 ? T : 4;
  const int nbg = row_blocks(ag.nrows, Tg, BS), npg = pad8(nbg);
  const int nbmet = row_blocks(amet.nrows, T, BS);
  DISPATCH_T3(T, {
    PFM_LAUNCH(h, "gm_first", (k_gm_first<T, R>), npt + npm, BS, at, Pt, it.B, it.Bp, BT, npt, nbt, am, Pr, h->eb,
               npm, nbm);
    if (energy)
      PFM_LAUNCH(h, "gm_second", (k_gm_second<T, R, true>), npg + npm, BS, ag, Pc, NN.nbr, simp, ic.gnorm, ic.gnf,
                 ic.gC, npg, nbg, am, Pr, h->eb, npm, nbm);
    else
      PFM_LAUNCH(h, "gm_second", (k_gm_second<T, R, false>), npg + npm, BS, ag, Pc, NN.nbr, simp, ic.gnorm, ic.gnf,
                 ic.gC, npg, nbg, am, Pr, h->eb, npm, nbm);
    const int g3 = npt + nbmet;
    if (energy && nd)
      PFM_LAUNCH(h, "gm_third", (k_gm_third<T, R, true, true>), g3, BS, at, Pt, row_of_tb, it.nl[0].rev, it.B, it.Bp,
                 BT, ndA, npt, nbt, amet, Pc, ic.gnorm, ic.gnf, ic.mrows, nbmet);
    else if (energy)
      PFM_LAUNCH(h, "gm_third", (k_gm_third<T, R, true, false>), g3, BS, at, Pt, row_of_tb, it.nl[0].rev, it.B, it.Bp,
                 BT, ndA, npt, nbt, amet, Pc, ic.gnorm, ic.gnf, ic.mrows, nbmet);
    else if (nd)
      PFM_LAUNCH(h, "gm_third", (k_gm_third<T, R, false, true>), g3, BS, at, Pt, row_of_tb, it.nl[0].rev, it.B, it.Bp,
                 BT, ndA, npt, nbt, amet, Pc, ic.gnorm, ic.gnf, ic.mrows, nbmet);
    else
      PFM_LAUNCH(h, "gm_third", (k_gm_third<T, R, false, false>), g3, BS, at, Pt, row_of_tb, it.nl[0].rev, it.B,
                 it.Bp, BT, ndA, npt, nbt, amet, Pc, ic.gnorm, ic.gnf, ic.mrows, nbmet);
  });
  return PFM_OK;
}

// The rjl rows and the graphene rows of the second launch write every atom's
// force first when they share no atom (h->gm.first_ok), no group changes are
// pending and their rows add up to N; the first launch writes no force and
// the third only adds.  The conditions under which gm_forces returns 1 are
// excluded here.
bool gm_first_writes(pfm_ctx* h) {
  const int T = h->lanes;
  if (!h->fuse || !h->gm.ok || !h->gm.first_ok || !h->changes.empty() || (T != 4 && T != 8 && T != 16)) return false;
  const Inter& ir = h->inter[h->gm.rjl];
  const Inter& it = h->inter[h->gm.tb];
  const Inter& ic = h->inter[h->gm.cos];
  if (ir.nl[0].nrows == 0 || it.nl[0].nrows == 0 || ic.nl[0].nrows == 0) return false;
  return (long)ir.nl[0].nrows + ic.nl[0].nrows == h->N;
}

template <class R>
static int forces_T(pfm_ctx* h, bool energy, bool first_writes) {
  bool written = !first_writes;  // has any pass written this step's forces yet
  if (h->fuse && h->gm.ok && (written || gm_first_writes(h))) {
    const int rc = gm_forces<R>(h, energy, !written);
    if (rc != 1) return rc;
  }
  if (!h->mp.planned) plan_multi(h);
  if (h->fuse && h->mp.ok && (written || multi_first_writes(h))) {
    const int rc = multi_forces<R>(h, energy, !written);
    if (rc != 1) return rc;
  }
  const int BS = std::min(h->block, PairBlock<R>::max);  // row passes (T lanes per row)
  const int BS1 = 256;           // one-lane-per-row passes (tb, cos_nn)
  for (size_t ii = 0; ii < h->inter.size(); ++ii) {
    Inter& it = h->inter[ii];
    switch (it.kind) {
      case PFM_RJL: {
        RowArgs a = row_args(h, it, 0, written ? 1 : 0);
        if (a.nrows == 0) break;
        RjlP P{it.p[0], it.p[1], it.p[2], it.p[3], it.p[4], it.p[5], it.p[6], 1.0 / (it.p[6] - it.p[5])};
        int T = pass_lanes(h, it);
        DISPATCH_T(T, {
          int nb = row_blocks(a.nrows, T, BS);
          PFM_LAUNCH(h, "rjl_density", (k_rjl_density<T, R>), pad8(nb), BS, a, P, h->eb, nb);
          if (energy) PFM_LAUNCH(h, "rjl_force", (k_rjl_force<T, R, true>), pad8(nb), BS, a, P, h->eb, nb);
          else PFM_LAUNCH(h, "rjl_force", (k_rjl_force<T, R, false>), pad8(nb), BS, a, P, h->eb, nb);
        });
        written = true;
        break;
      }
      case PFM_LJ: {
        LjP P{4.0 * it.p[0], it.p[1], it.p[2], it.p[3], 1.0 / (it.p[3] - it.p[2])};
        for (int k = 0; k < 2; ++k) {
          RowArgs a = row_args(h, it, k, written ? 1 : 0);
          if (a.nrows == 0) continue;
          bool E = energy && k == 0;
          int T = pass_lanes(h, it);
          DISPATCH_T(T, {
            int nb = row_blocks(a.nrows, T, BS);
            if (E) PFM_LAUNCH(h, "lj", (k_lj<T, R, true>), pad8(nb), BS, a, P, nb);
            else PFM_LAUNCH(h, "lj", (k_lj<T, R, false>), pad8(nb), BS, a, P, nb);
          });
          written = true;
        }
        break;
      }
      case PFM_LJ1G: {
        RowArgs a = row_args(h, it, 0, written ? 1 : 0);
        if (a.nrows == 0) break;
        double eps = it.p[0], sig = it.p[1], s6 = sig * sig * sig * sig * sig * sig;
        Lj1gP P{4.0 * eps * s6, 4.0 * eps * s6 * s6, 6.0 * 4.0 * eps * s6, 12.0 * 4.0 * eps * s6 * s6,
                it.p[2], it.p[3], 1.0 / (it.p[3] - it.p[2])};
        int T = pass_lanes(h, it);
        DISPATCH_T(T, {
          int nb = row_blocks(a.nrows, T, BS);
          if (energy) PFM_LAUNCH(h, "lj1g", (k_lj1g<T, R, true>), pad8(nb), BS, a, P, nb);
          else PFM_LAUNCH(h, "lj1g", (k_lj1g<T, R, false>), pad8(nb), BS, a, P, nb);
        });
        written = true;
        break;
      }
      case PFM_TB: {
        RowArgs a = row_args(h, it, 0, written ? 1 : 0);
        if (a.nrows == 0) break;
        TbP P{it.p[0], it.p[1], it.p[2], it.p[3], it.p[4], it.p[5], it.p[6] * it.p[6], it.p[7] * it.p[7],
              it.p[8], it.p[9], 1.0 / (it.p[9] - it.p[8])};
        int nb = row_blocks(a.nrows, TB_T, BS1);
        double* BT = it.Bp + (size_t)it.nl[0].nrows * it.nl[0].stride;
        PFM_LAUNCH(h, "tb_bond", (k_tb_bond<R>), pad8(nb), BS1, a, P, it.B, it.Bp, BT, nb);
        const int* row_of = h->grp.row_of[it.nl[0].g1];
        if (energy) PFM_LAUNCH(h, "tb_force", (k_tb_force<R, true>), pad8(nb), BS1, a, P, row_of, it.nl[0].rev, it.B, it.Bp, BT, nb);
        else PFM_LAUNCH(h, "tb_force", (k_tb_force<R, false>), pad8(nb), BS1, a, P, row_of, it.nl[0].rev, it.B, it.Bp, BT, nb);
        written = true;
        break;
      }
      case PFM_REBOSC: {
        int rc = launch_rebo(h, it, energy, written ? 1 : 0);
        if (rc) return rc;
        written = true;
        break;
      }
      case PFM_SUR: {
        int rc = launch_sur_forces(h, it, written ? 1 : 0);
        if (rc) {
          h->err = "error: surrogate parameters not set";
          return rc;
        }
        written = true;
        break;
      }
      case PFM_LJC:
      case PFM_MORSEC: {
        bool morse = it.kind == PFM_MORSEC;
        bool simplified = (int)it.p[morse ? 6 : 5];
        CosP P;
        if (morse) {
          P = CosP{it.p[0], 0.0, it.p[2], it.p[1], it.p[3], it.p[4], it.p[5], 1.0 / (it.p[5] - it.p[4]),
                   2.0 * it.p[0] * it.p[3], 1};
        } else {
          P = CosP{4.0 * it.p[0], it.p[1], 0.0, 0.0, it.p[2], it.p[3], it.p[4], 1.0 / (it.p[4] - it.p[3]),
                   4.0 * it.p[0] * it.p[2], 0};
        }
        ListDev& NN = it.nl[2];
        RowArgs a = row_args(h, it, 0, written ? 1 : 0);
        int T = pass_lanes(h, it);
        int simp = simplified ? 1 : 0;
        if (a.nrows > 0) {
          // graphene rows hold about 40 metal entries: at most four lanes each
          const int Tg = T > 4 ? 4 : T;
          DISPATCH_T(Tg, {
            int nb = row_blocks(a.nrows, T, BS);
            if (energy) PFM_LAUNCH(h, "cos_graphene", (k_cos_graphene<T, R, true>), pad8(nb), BS, a, P, NN.nbr, simp, it.gnorm, it.gnf, it.gC, nb);
            else PFM_LAUNCH(h, "cos_graphene", (k_cos_graphene<T, R, false>), pad8(nb), BS, a, P, NN.nbr, simp, it.gnorm, it.gnf, it.gC, nb);
          });
          written = true;
        }
        RowArgs c = row_args(h, it, 1, written ? 1 : 0);
        if (c.nrows > 0 && !c.accumulate) {
          // the metal pass skips rows without entries, so it never writes first
          h->err = "error: the ljc/morsec metal pass would write the first forces of its step";
          return PFM_E_INVALID;
        }
        const bool nn = !simplified && NN.nrows > 0;
        if (nn && c.nrows > 0 && it.rows_disjoint && BS == BS1) {
          RowArgs b = row_args(h, it, 2, written ? 1 : 0);
          int nbn = row_blocks(b.nrows, 1, BS1);
          DISPATCH_T(T, {
            int nb = nbn + row_blocks(c.nrows, T, BS1);
            PFM_LAUNCH(h, "cos_nn_metal", (k_cos_nn_metal<T, R>), pad8(nb), BS1, b, c, P, h->grp.row_of[NN.g1],
                       NN.rev, it.gC, it.gnorm, it.gnf, it.mrows, nbn, nb);
          });
          written = true;
          break;
        }
        if (nn) {
          RowArgs b = row_args(h, it, 2, written ? 1 : 0);
          int nb = row_blocks(b.nrows, 1, BS1);
          PFM_LAUNCH(h, "cos_nn", k_cos_nn, pad8(nb), BS1, b, h->grp.row_of[NN.g1], NN.rev, it.gC, nb);
          written = true;
        }
        c.accumulate = written ? 1 : 0;
        if (c.nrows > 0) {
          DISPATCH_T(T, {
            int nb = row_blocks(c.nrows, T, BS);
            PFM_LAUNCH(h, "cos_metal", (k_cos_metal<T, R>), pad8(nb), BS, c, P, it.gnorm, it.gnf, it.mrows, nb);
          });
          written = true;
        }
        break;
      }
    }
  }
  return PFM_OK;
}

int launch_forces(pfm_ctx* h, bool energy, bool first_writes) {
  if (h->precision == PFM_MIXED) return forces_T<float>(h, energy, first_writes);
  return forces_T<double>(h, energy, first_writes);
}
