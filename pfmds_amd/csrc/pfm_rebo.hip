// pfm_rebo.hip -- rebosc, the REBO solid-carbon potential (REBOsolidcarbon.f90)
// on gfx950.
//
// The reference evaluates only the energy (REBOsc_energy, :27-92) and gets
// forces by central differences, dx = 1e-6 A, of the energy of a list
// truncated to three hops around each atom (calculate_forces_numerically,
// md_interactions.f90:273-310).  The atoms whose position moves a bond's
// energy are its two ends and their neighbours, so that difference equals the
// difference of the energies of the bonds with an end within one hop of the
// moved atom; k_rebo_force forms exactly those, with the same dx, in FP64
// (both precisions): eight lanes per atom, lanes 2k and 2k+1 the -dx and +dx
// energies along axis k.  Bond energies are recomputed from the positions
// (no stored distances or bond orders), each bond counted once from its end
// with the lower group-local index, as the reference's energy loop does.
#include <algorithm>

#include "pfm_internal.h"

struct ReboP {
  double A, Q, al, B[3], be[3], T, g[6], R1, R2, inv_w;
};

struct ReboArgs {
  int nrows, stride;
  const int* rows;    // atom of every row
  const int* row_of;  // row of every atom (-1 outside the group)
  const int* nnum;
  const int* nbr;
  const int* rev;     // slot of the row atom in the neighbour's row
  const int* loc;     // group-local index by original atom
  const int* orig;
  const double4* pos;
  double3 L, iL;
  const int* rep;
  const RepDev* rdev;
  double4* frc;
  int accumulate;
  double* e_row;
};

// x_b - x_a (minimum image) with atom s displaced by d
__device__ __forceinline__ double3 rdr(const ReboArgs& A, int a, int b, int s, double3 d) {
  double3 L = A.L, iL = A.iL;
  if (A.rep) {
    const RepDev& R = A.rdev[A.rep[a]];
    L = R.L;
    iL = R.iL;
  }
  double x, y, z;
  fast_dist2(A.pos[a], A.pos[b], L, iL, x, y, z);
  if (b == s) { x += d.x; y += d.y; z += d.z; }
  if (a == s) { x -= d.x; y -= d.y; z -= d.z; }
  return make_double3(x, y, z);
}

__device__ __forceinline__ double dot3(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ double rebo_fc(const ReboP& P, double r) {
  double f, df;
  cos_cutoff<double>(r, P.R1, P.R2, P.inv_w, f, df);
  return f;
}

// bsp of bond (a, slot p) whose separation is dp (REBOsolidcarbon.f90:44-46,63)
__device__ double rebo_bsp(const ReboArgs& A, const ReboP& P, int ra, int a, int p, double3 dp, double rp, int s,
                           double3 d) {
  const int n = A.nnum[ra];
  const int* nb = A.nbr + (size_t)ra * A.stride;
  double sum = 0.0;
  for (int q = 0; q < n; ++q) {
    if (q == p) continue;
    double3 dq = rdr(A, a, nb[q], s, d);
    double rq = sqrt(dot3(dq, dq));
    if (rq < P.R2) {
      double c = dot3(dp, dq) / (rp * rq);
      double poly = P.g[0] + c * (P.g[1] + c * (P.g[2] + c * (P.g[3] + c * (P.g[4] + c * P.g[5]))));
      sum += rebo_fc(P, rq) * poly;
    }
  }
  return 1.0 / sqrt(1.0 + sum);
}

// energy of bond (a, slot p) counted from a (REBOsolidcarbon.f90:70-89)
__device__ double rebo_bond(const ReboArgs& A, const ReboP& P, int ra, int a, int p, int s, double3 d) {
  const int* nb = A.nbr + (size_t)ra * A.stride;
  const int j = nb[p];
  double3 dp = rdr(A, a, j, s, d);
  double rp = sqrt(dot3(dp, dp));
  if (!(rp < P.R2)) return 0.0;
  double b1 = rebo_bsp(A, P, ra, a, p, dp, rp, s, d);
  const int rj = A.row_of[j];
  const int qj = A.rev[(size_t)ra * A.stride + p];
  double b2 = 0.0;
  if (rj >= 0 && qj >= 0) {
    double3 dj = rdr(A, j, a, s, d);
    b2 = rebo_bsp(A, P, rj, j, qj, dj, sqrt(dot3(dj, dj)), s, d);
  }
  // dihedral term (REBOsolidcarbon.f90:47-58)
  double bdh = 0.0;
  if (rj >= 0) {
    const int n = A.nnum[ra], nj = A.nnum[rj];
    const int* nbj = A.nbr + (size_t)rj * A.stride;
    const double aa = rp * rp;
    for (int q = 0; q < n; ++q) {
      if (q == p) continue;
      double3 dq = rdr(A, a, nb[q], s, d);
      double rq = sqrt(dot3(dq, dq));
      if (!(rq < P.R2)) continue;
      double fq = rebo_fc(P, rq), ab = dot3(dp, dq);
      for (int l = 0; l < nj; ++l) {
        if (nbj[l] == a) continue;
        double3 dl = rdr(A, j, nbj[l], s, d);
        double rl = sqrt(dot3(dl, dl));
        if (!(rl < P.R2)) continue;
        double ac = dot3(dp, dl), bc = dot3(dq, dl);
        double x = aa * bc - ab * ac;
        bdh += fq * rebo_fc(P, rl) * (1.0 - x * x / (aa * rq * rq - ab * ab) / (aa * rl * rl - ac * ac));
      }
    }
  }
  bdh *= P.T;
  double att = P.B[0] * exp(-P.be[0] * rp) + P.B[1] * exp(-P.be[1] * rp) + P.B[2] * exp(-P.be[2] * rp);
  return rebo_fc(P, rp) * ((1.0 + P.Q / rp) * P.A * exp(-P.al * rp) - ((b1 + b2) / 2.0 + bdh) * att);
}

__device__ __forceinline__ int rebo_loc(const ReboArgs& A, int atom) { return A.loc[A.orig[atom]]; }

// per-row energy: the bonds the row's atom counts (REBOsolidcarbon.f90:66-89)
__global__ void __launch_bounds__(256) k_rebo_energy(ReboArgs A, ReboP P) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.nrows) return;
  const int a = A.rows[r];
  const int la = rebo_loc(A, a);
  const int* nb = A.nbr + (size_t)r * A.stride;
  double e = 0.0;
  for (int p = 0; p < A.nnum[r]; ++p)
    if (rebo_loc(A, nb[p]) > la) e += rebo_bond(A, P, r, a, p, -1, make_double3(0, 0, 0));
  A.e_row[r] = e;
}

#define REBO_MAXNB 32

// forces by central differences: lane 2k+h of an atom's eight lanes gets the
// energy of the bonds around the atom displaced by (2h-1) dx along axis k
__global__ void __launch_bounds__(256) k_rebo_force(ReboArgs A, ReboP P, int nb_real) {
  int lb = xcd_block(nb_real);
  if (lb < 0) return;
  int gt = lb * blockDim.x + threadIdx.x;
  int r = gt >> 3, t = threadIdx.x & 7;
  if (r >= A.nrows) return;
  const int s = A.rows[r];
  const double dx = 1e-6;
  double e = 0.0;
  if (t < 6) {
    const int k = t >> 1;
    const double h = (t & 1) ? dx : -dx;
    const double3 d = make_double3(k == 0 ? h : 0.0, k == 1 ? h : 0.0, k == 2 ? h : 0.0);
    // the moved atom and its bonded neighbours (unmoved distances)
    int hood[REBO_MAXNB + 1], nh = 0;
    hood[nh++] = s;
    const int* nbs = A.nbr + (size_t)r * A.stride;
    for (int p = 0; p < A.nnum[r] && nh <= REBO_MAXNB; ++p) {
      double3 dp = rdr(A, s, nbs[p], -1, d);
      if (dot3(dp, dp) < P.R2 * P.R2) hood[nh++] = nbs[p];
    }
    for (int m = 0; m < nh; ++m) {
      const int a = hood[m], ra = A.row_of[a], la = rebo_loc(A, a);
      const int* nb = A.nbr + (size_t)ra * A.stride;
      for (int p = 0; p < A.nnum[ra]; ++p) {
        const int b = nb[p], lb2 = rebo_loc(A, b);
        if (la < lb2) {
          e += rebo_bond(A, P, ra, a, p, s, d);
        } else {
          bool in = false;
          for (int u = 0; u < nh; ++u) in = in || hood[u] == b;
          if (!in) {
            const int rb = A.row_of[b], q = A.rev[(size_t)ra * A.stride + p];
            if (rb >= 0 && q >= 0) e += rebo_bond(A, P, rb, b, q, s, d);
          }
        }
      }
    }
  }
  // (e(-dx) - e(+dx)) / (2 dx) along each axis: lanes 2k and 2k+1
  double other = __shfl_xor(e, 1, 8);
  double f = (e - other) / (2.0 * dx);  // -dE/dx on the lanes 2k (e = e(-dx), other = e(+dx))
  double fy = __shfl(f, (threadIdx.x & ~7) + 2);
  double fz = __shfl(f, (threadIdx.x & ~7) + 4);
  if (t == 0) {
    double4 F = A.accumulate ? A.frc[s] : make_double4(0, 0, 0, 0);
    F.x += f;
    F.y += fy;
    F.z += fz;
    A.frc[s] = F;
  }
}

static ReboArgs rebo_args(pfm_ctx* h, Inter& it, int accumulate) {
  ListDev& L = it.nl[0];
  ReboArgs a;
  a.nrows = L.nrows;
  a.stride = L.stride;
  a.rows = h->grp.rows[L.g1];
  a.row_of = h->grp.row_of[L.g1];
  a.nnum = L.nnum;
  a.nbr = L.nbr;
  a.rev = L.rev;
  a.loc = h->grp.loc[L.g1];
  a.orig = h->orig;
  a.pos = h->pos4;
  a.L = make_double3(h->box[0], h->box[1], h->box[2]);
  a.iL = make_double3(1.0 / h->box[0], 1.0 / h->box[1], 1.0 / h->box[2]);
  a.rep = h->rep;
  a.rdev = h->rdev;
  a.frc = h->frc4;
  a.accumulate = accumulate;
  a.e_row = it.e_row;
  return a;
}

int launch_rebo(pfm_ctx* h, Inter& it, bool energy, int accumulate) {
  ReboArgs a = rebo_args(h, it, accumulate);
  if (a.nrows == 0) return PFM_OK;
  const double* p = it.p;
  ReboP P;
  P.A = p[0]; P.Q = p[1]; P.al = p[2];
  for (int k = 0; k < 3; ++k) { P.B[k] = p[3 + k]; P.be[k] = p[6 + k]; }
  P.T = p[9];
  for (int k = 0; k < 6; ++k) P.g[k] = p[10 + k];
  P.R1 = p[16]; P.R2 = p[17];
  P.inv_w = 1.0 / (P.R2 - P.R1);
  if (energy) PFM_LAUNCH(h, "rebo_energy", k_rebo_energy, (a.nrows + 255) / 256, 256, a, P);
  int nb = (int)(((long)a.nrows * 8 + 255) / 256);
  PFM_LAUNCH(h, "rebo_force", k_rebo_force, pad8(nb), 256, a, P, nb);
  return PFM_OK;
}
