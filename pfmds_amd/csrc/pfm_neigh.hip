// pfm_neigh.hip -- cell binning, Morton re-sort of the atoms and neighbour-list
// construction on gfx950.
//
// Replaces find_neighbours (md_neighbours.f90:56-100, brute force over all
// N1*N2 pairs), converce_neighbour_list (md_neighbours.f90:128-160, a serial
// transpose) and find_gr_nearest_neighbors (graphenenorm.f90:8-36).  The
// membership test is the reference's own: j != i, j in group 2, and
// dr2 < r_cut**2 with find_distance's arithmetic (min_image_dist2), so the
// lists are identical to the reference's as sets.  Lists are rebuilt on the
// reference's schedule (mod(md_step, update_period) == 0); between rebuilds
// nothing is refreshed because the force passes recompute dr from positions.
#include "pfm_internal.h"

// ------------------------------------------------------------------ binning

__device__ __forceinline__ int cell_coord(double x, double inv_len, int nc) {
  int c = (int)floor(x * inv_len);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_fill_int(int n, int* __restrict__ x, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// ctr[key] += 1 for every valid lane of the wavefront, with one atomic per run
// of equal keys in consecutive lanes (atoms arrive nearly in cell order, so
// runs are long); returns the lane's previous count (its slot).  Invalid
// lanes (past the end of the array) are the top lanes of the last wavefront.
__device__ __forceinline__ int run_atomic_inc(int* ctr, int key, bool valid) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1);
  const bool head = valid && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const int nvalid = __popcll(__ballot(valid));
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
  const int start = valid ? 63 - __clzll(heads & upto) : lane;
  const unsigned long long above = heads & ~upto;
  const int next = above ? __ffsll(above) - 1 : 64;
  int base = 0;
  if (head) base = atomicAdd(&ctr[key], (next < nvalid ? next : nvalid) - lane);
  base = __shfl(base, start);
  return base + (lane - start);
}

__global__ void k_cell_keys(int N, const double4* __restrict__ pos, const int* __restrict__ rep,
                            const RepDev* __restrict__ rdev, const int* __restrict__ cell_rank,
                            int* __restrict__ atom_rank, int* __restrict__ cell_count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < N;
  int r = -1;
  if (valid) {
    const RepDev& R = rdev[rep ? rep[i] : 0];
    double4 p = pos[i];
    int cx = cell_coord(p.x, R.inv_clen.x, R.nc[0]);
    int cy = cell_coord(p.y, R.inv_clen.y, R.nc[1]);
    int cz = cell_coord(p.z, R.inv_clen.z, R.nc[2]);
    r = cell_rank[R.cell_base + (cz * R.nc[1] + cy) * R.nc[0] + cx];
    atom_rank[i] = r;
  }
  run_atomic_inc(cell_count, r, valid);
}

// exclusive scan of n ints by one workgroup of 1024 threads (small n: the
// cell counts, and the block sums of scan_excl below)
__device__ __forceinline__ void scan_excl_wg(int n, const int* __restrict__ in, int* __restrict__ out,
                                             int* __restrict__ total) {
  __shared__ int s[1024];
  int t = threadIdx.x;
  int chunk = (n + 1023) / 1024;
  int b = t * chunk, e = min(n, b + chunk);
  int sum = 0;
  for (int k = b; k < e; ++k) sum += in[k];
  s[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = s[t] - sum;
  for (int k = b; k < e; ++k) {
    int v = in[k];
    out[k] = run;
    run += v;
  }
  if (t == 1023 && total) *total = s[1023];
}

__global__ void __launch_bounds__(1024) k_scan_excl(int n, const int* __restrict__ in, int* __restrict__ out,
                                                    int* __restrict__ total) {
  scan_excl_wg(n, in, out, total);
}

#define SCAN_B 1024  // items per workgroup of the multi-block scan (256 threads x 4)

__device__ __forceinline__ int wg_inclusive_scan(int v, int* sh) {
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    int w = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += w;
    __syncthreads();
  }
  int r = sh[t];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_scan_blocksum(int n, const int* __restrict__ in, int* __restrict__ bsum) {
  __shared__ int sh[256];
  int base = blockIdx.x * SCAN_B + threadIdx.x * 4, v = 0;
  for (int k = 0; k < 4; ++k)
    if (base + k < n) v += in[base + k];
  int inc = wg_inclusive_scan(v, sh);
  if (threadIdx.x == 255) bsum[blockIdx.x] = inc;
}

__global__ void __launch_bounds__(256) k_scan_apply(int n, const int* __restrict__ in, const int* __restrict__ boff,
                                                    int* __restrict__ out) {
  __shared__ int sh[256];
  int base = blockIdx.x * SCAN_B + threadIdx.x * 4;
  int x[4], v = 0;
  for (int k = 0; k < 4; ++k) {
    x[k] = base + k < n ? in[base + k] : 0;
    v += x[k];
  }
  int run = wg_inclusive_scan(v, sh) - v + boff[blockIdx.x];
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += x[k];
  }
}

// exclusive scan of n ints; tmp holds ceil(n/1024) block sums and their scan
void scan_excl(pfm_ctx* h, int n, const int* in, int* out, int* tmp) {
  int nb = (n + SCAN_B - 1) / SCAN_B;
  if (n <= 16 * 1024) {
    // one workgroup, up to 16 items a thread: one launch instead of three
    PFM_LAUNCH(h, "scan", k_scan_excl, 1, 1024, n, in, out, (int*)nullptr);
    return;
  }
  PFM_LAUNCH(h, "scan", k_scan_blocksum, nb, 256, n, in, tmp);
  PFM_LAUNCH(h, "scan", k_scan_excl, 1, 1024, nb, tmp, tmp + nb, (int*)nullptr);
  PFM_LAUNCH(h, "scan", k_scan_apply, nb, 256, n, in, tmp + nb, out);
}

__global__ void k_scatter(int N, const int* __restrict__ atom_rank, const int* __restrict__ cell_start,
                          int* __restrict__ cell_fill, int* __restrict__ perm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < N;
  const int r = valid ? atom_rank[i] : -1;
  const int slot = run_atomic_inc(cell_fill, r, valid);
  if (valid) perm[cell_start[r] + slot] = i;
}

// the per-atom arrays in their current order (0) and in cell order (1)
struct PermArgs {
  const double4 *p0, *v0, *f0;
  const uint32_t* m0;
  const int *o0, *r0;
  double4 *p1, *v1, *f1;
  uint32_t* m1;
  int *o1, *r1;
  uint4* q1;  // fixed-point copy of the new positions (mixed path), or null
  const RepDev* rdev;
};

// order the atoms of each cell by original index, so the layout (and every
// sum over it) does not depend on the order of the scatter's atomics: one
// 64-lane workgroup per cell, each lane ranks its entries against the cell.
// PERM: every atom's data moves to its new place (the re-sort); otherwise
// the order goes to perm_out (binning without moving the atoms).  Also the
// cell's count for the searches, the union of its group bits (cell_grp,
// which lets a list search skip tiles that hold no candidate), and the
// binning's counters cleared for the next binning.
template <bool PERM>
__global__ void __launch_bounds__(64) k_cell_sort(int ncell, const int* __restrict__ cell_start,
                                                  int* __restrict__ cell_tmp, int* __restrict__ cell_fill,
                                                  int* __restrict__ cell_count, const int* __restrict__ perm_in,
                                                  int* __restrict__ perm_out, const int* __restrict__ orig,
                                                  const uint32_t* __restrict__ mask, uint32_t* __restrict__ cell_grp,
                                                  PermArgs P) {
  __shared__ int key[512];
  int c = blockIdx.x;
  if (c >= ncell) return;
  int s = cell_start[c], n = cell_tmp[c];
  bool lds = n <= 512;
  uint32_t g = 0u;
  if (lds)
    for (int k = threadIdx.x; k < n; k += 64) key[k] = orig[perm_in[s + k]];
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += 64) {
    int v = perm_in[s + k];
    g |= mask[v];
    int kk = lds ? key[k] : orig[v];
    int rank = 0;
    for (int m = 0; m < n; ++m) rank += ((lds ? key[m] : orig[perm_in[s + m]]) < kk) ? 1 : 0;
    if constexpr (PERM) {
      const int d = s + rank;
      const double4 p = P.p0[v];
      P.p1[d] = p;
      int r = 0;
      if (P.r0) { r = P.r0[v]; P.r1[d] = r; }
      if (P.q1) P.q1[d] = quantize(p, P.rdev[r].qs, 0u);
      P.v1[d] = P.v0[v];
      P.f1[d] = P.f0[v];
      P.m1[d] = P.m0[v];
      P.o1[d] = P.o0[v];
    } else {
      perm_out[s + rank] = v;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) g |= (uint32_t)__shfl_xor((int)g, o);
  if (threadIdx.x == 0) {
    cell_grp[c] = g;
    cell_count[c] = n;
    cell_tmp[c] = 0;
    cell_fill[c] = 0;
  }
}

// Bin atoms into cells and re-order every per-atom array in Morton cell
// order.  The binning counters (cell_tmp, cell_fill) are zero on entry: the
// allocation clears them and k_cell_sort clears them again for the next call.
int launch_sort_and_bin(pfm_ctx* h, int step, bool permute) {
  (void)step;
  int N = h->N, nb = (N + 255) / 256;
  PFM_LAUNCH(h, "cell_keys", k_cell_keys, nb, 256, N, h->pos4, h->rep, h->rdev, h->cell_rank, h->atom_rank,
             h->cell_tmp);
  scan_excl(h, h->ncell, h->cell_tmp, h->cell_start, h->scan_tmp);
  PFM_LAUNCH(h, "scatter", k_scatter, nb, 256, N, h->atom_rank, h->cell_start, h->cell_fill, h->scan_tmp + N);
  PermArgs P{h->pos4, h->vel4, h->frc4, h->mask, h->orig, h->rep, h->pos4b, h->vel4b, h->frc4b, h->maskb,
             h->origb, h->repb, h->qpos, h->rdev};
  if (!permute) {
    // lists that are not due keep their internal indices: the search reads
    // the atoms of every cell through the cell-ordered index
    PFM_LAUNCH(h, "cell_sort", k_cell_sort<false>, h->ncell, 64, h->ncell, h->cell_start, h->cell_tmp,
               h->cell_fill, h->cell_count, h->scan_tmp + N, h->perm, h->orig, h->mask, h->cell_grp, P);
    h->bin = h->perm;
    return PFM_OK;
  }
  PFM_LAUNCH(h, "cell_sort", k_cell_sort<true>, h->ncell, 64, h->ncell, h->cell_start, h->cell_tmp, h->cell_fill,
             h->cell_count, h->scan_tmp + N, h->perm, h->orig, h->mask, h->cell_grp, P);
  std::swap(h->pos4, h->pos4b);
  std::swap(h->vel4, h->vel4b);
  std::swap(h->frc4, h->frc4b);
  std::swap(h->mask, h->maskb);
  std::swap(h->orig, h->origb);
  std::swap(h->rep, h->repb);
  h->bin = nullptr;  // atoms are in cell order
  launch_group_rows(h);
  return PFM_OK;
}

// --------------------------------------------------------------- group rows
//
// The rows of up to GR_MAX groups in three launches whatever their number:
// per 256-atom block and group the member count (wave ballots), one
// workgroup per group scanning its block counts, then every member's row
// from its block's offset, its wave's offset and its rank in the wave.  Rows
// come out in internal order, as a stable compaction gives them.

#define GR_MAX 8
struct GroupRowsArgs {
  int ng, nb;
  uint32_t bit[GR_MAX];
  int* rows[GR_MAX];
  int* row_of[GR_MAX];
};

__global__ void __launch_bounds__(256) k_grp_count(int N, const uint32_t* __restrict__ mask, GroupRowsArgs G,
                                                   int* __restrict__ cnt) {
  __shared__ int wc[GR_MAX][4];
  const int i = blockIdx.x * 256 + threadIdx.x, w = threadIdx.x >> 6;
  const uint32_t m = i < N ? mask[i] : 0u;
  for (int q = 0; q < G.ng; ++q) {
    const unsigned long long b = __ballot((m & G.bit[q]) != 0);
    if ((threadIdx.x & 63) == 0) wc[q][w] = __popcll(b);
  }
  __syncthreads();
  if (threadIdx.x < G.ng) {
    const int q = threadIdx.x;
    cnt[q * G.nb + blockIdx.x] = wc[q][0] + wc[q][1] + wc[q][2] + wc[q][3];
  }
}

__global__ void __launch_bounds__(1024) k_grp_scan(int nb, int* __restrict__ cnt, int* __restrict__ off) {
  const int q = blockIdx.x;
  scan_excl_wg(nb, cnt + (size_t)q * nb, off + (size_t)q * nb, nullptr);
}

__global__ void __launch_bounds__(256) k_grp_write(int N, const uint32_t* __restrict__ mask, GroupRowsArgs G,
                                                   const int* __restrict__ off) {
  __shared__ int wc[GR_MAX][4];
  const int i = blockIdx.x * 256 + threadIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t m = i < N ? mask[i] : 0u;
  unsigned long long bal[GR_MAX];
  for (int q = 0; q < G.ng; ++q) {
    bal[q] = __ballot((m & G.bit[q]) != 0);
    if (lane == 0) wc[q][w] = __popcll(bal[q]);
  }
  __syncthreads();
  if (i >= N) return;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int q = 0; q < G.ng; ++q) {
    if (m & G.bit[q]) {
      int r = off[q * G.nb + blockIdx.x] + __popcll(bal[q] & below);
      for (int v = 0; v < w; ++v) r += wc[q][v];
      G.rows[q][r] = i;
      G.row_of[q][i] = r;
    } else {
      G.row_of[q][i] = -1;
    }
  }
}

__global__ void k_identity_rows(int N, int* __restrict__ rows, int* __restrict__ row_of) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) { rows[i] = i; row_of[i] = i; }
}

// rows of every group that some list or role uses: members in internal order
void launch_group_rows(pfm_ctx* h) {
  int N = h->N, nb = (N + 255) / 256;
  GroupRowsArgs G{};
  G.nb = nb;
  auto flush = [&]() {
    if (!G.ng) return;
    int* cnt = h->scan_tmp;
    int* off = h->scan_tmp + (size_t)GR_MAX * nb;
    PFM_LAUNCH(h, "group_rows", k_grp_count, nb, 256, N, h->mask, G, cnt);
    PFM_LAUNCH(h, "group_rows", k_grp_scan, G.ng, 1024, nb, cnt, off);
    PFM_LAUNCH(h, "group_rows", k_grp_write, nb, 256, N, h->mask, G, off);
    G.ng = 0;
  };
  for (int g = 1; g <= PFM_MAXG; ++g) {
    if (!h->grp.used[g] || h->grp.size[g] == 0) continue;
    // members (a growing group's spare bit), not only the active atoms
    uint32_t bit = 1u << ((h->grp.shadow[g] ? h->grp.shadow[g] : g) - 1);
    if (h->grp.size[g] == N) {
      // every atom is a member: row r is internal atom r whatever the order
      if (!h->grp.ident[g]) PFM_LAUNCH(h, "group_rows", k_identity_rows, nb, 256, N, h->grp.rows[g], h->grp.row_of[g]);
      h->grp.ident[g] = true;
      continue;
    }
    G.bit[G.ng] = bit;
    G.rows[G.ng] = h->grp.rows[g];
    G.row_of[G.ng] = h->grp.row_of[g];
    if (++G.ng == GR_MAX) flush();
  }
  flush();
}

// ------------------------------------------------------------ list building

struct SearchArgs {
  int nrows;
  const int* rows;
  uint32_t g1bit;  // active row atoms; the rows of a growing group hold inactive members with no entries
  uint32_t g1mem;  // members of group 1 (a growing group's spare bit): the atoms that have rows
  uint32_t g2bit;
  const double4* pos;
  const uint32_t* mask;
  const int* orig;
  const RepDev* rdev;   // replica geometry
  const int* rep;       // replica of every internal atom (null: one replica)
  const int* cell_rep;  // replica of every cell (null: one replica)
  const int* bin;       // atoms of the cells in cell order (null: atoms are stored in cell order)
  double rc2, rc2_lo, rc2_hi;
  const int* cell_rank;
  const int* cell_start;
  const int* cell_count;
  const uint32_t* cell_grp;  // union of the group bits of each cell's atoms
  int nmax, stride;
  int* nnum;
  int* nbr;
  DevErr* err;
  int step, err_code;
  int exact;  // require exactly nmax entries (graphene nearest neighbours)
  // a second list with the same rows, cutoff and rebuild steps searched in
  // the same pass (k_build_tile_wave<T, CMP, true>): g2bit is then the union
  // of the two candidate groups, g2bit1 / g2bit2 the groups of the two lists
  uint32_t g2bit1, g2bit2;
  int nmax2, stride2;
  int* nnum2;
  int* nbr2;
  // a second pair in the same pass (k_build_tile_wave<T, CMP, true, true>):
  // rows of another group (side b, disjoint from group 1), whose two lists
  // take the candidates of g2bit1 (b1) and of g2bit2 (b2); g1mem is then
  // the union of the two row groups' member bits
  uint32_t g1bit_b, g1mem_b;
  const int* row_of_b;
  int nmax_b1, stride_b1, nmax_b2, stride_b2;
  int *nnum_b1, *nbr_b1, *nnum_b2, *nbr_b2;
};

// T lanes per row atom scan the 27 (or fewer, for boxes under three cells
// wide) surrounding cells; wave ballots compact the accepted candidates.
template <int T>
__global__ void __launch_bounds__(256) k_build(SearchArgs a, int nblocks_real) {
  int lb = xcd_block(nblocks_real);
  if (lb < 0) return;
  int gt = lb * blockDim.x + threadIdx.x;
  int row = gt / T, t = threadIdx.x & (T - 1);
  if (row >= a.nrows) return;
  int lane = threadIdx.x & 63;
  int gshift = lane & ~(T - 1);
  unsigned long long gmask = (T == 64) ? ~0ull : (((1ull << T) - 1ull) << gshift);
  unsigned long long lower = (1ull << lane) - 1ull;
  int i = a.rows[row];
  const RepDev R = a.rdev[a.rep ? a.rep[i] : 0];
  if (R.tile) return;  // searched by k_build_tile
  if (!(a.mask[i] & a.g1bit)) {
    // find_neighbours runs over the group's active atoms only (md_neighbours.f90:67)
    if (t == 0) a.nnum[row] = 0;
    return;
  }
  double4 pi = a.pos[i];
  int c[3];
  c[0] = cell_coord(pi.x, R.inv_clen.x, R.nc[0]);
  c[1] = cell_coord(pi.y, R.inv_clen.y, R.nc[1]);
  c[2] = cell_coord(pi.z, R.inv_clen.z, R.nc[2]);
  int lo[3], cnt[3];
  for (int d = 0; d < 3; ++d) {
    if (R.nc[d] >= 3) { lo[d] = c[d] - 1; cnt[d] = 3; }
    else { lo[d] = 0; cnt[d] = R.nc[d]; }
  }
  int count = 0;
  int* out = a.nbr + (size_t)row * a.stride;
  for (int oz = 0; oz < cnt[2]; ++oz) {
    int z = lo[2] + oz; z = z < 0 ? z + R.nc[2] : (z >= R.nc[2] ? z - R.nc[2] : z);
    for (int oy = 0; oy < cnt[1]; ++oy) {
      int y = lo[1] + oy; y = y < 0 ? y + R.nc[1] : (y >= R.nc[1] ? y - R.nc[1] : y);
      for (int ox = 0; ox < cnt[0]; ++ox) {
        int x = lo[0] + ox; x = x < 0 ? x + R.nc[0] : (x >= R.nc[0] ? x - R.nc[0] : x);
        int r = a.cell_rank[R.cell_base + (z * R.nc[1] + y) * R.nc[0] + x];
        int s = a.cell_start[r], n = a.cell_count[r];
        for (int base = 0; base < n; base += T) {
          int k = base + t;
          bool pred = false;
          int j = 0;
          if (k < n) {
            j = a.bin ? a.bin[s + k] : s + k;
            if (j != i && (a.mask[j] & a.g2bit)) {
              double4 pj = a.pos[j];
              double ax = pj.x - pi.x, ay = pj.y - pi.y, az = pj.z - pi.z;
              ax = fma(-R.L.x, rint(ax * R.iL.x), ax);
              ay = fma(-R.L.y, rint(ay * R.iL.y), ay);
              az = fma(-R.L.z, rint(az * R.iL.z), az);
              double d2 = fma(ax, ax, fma(ay, ay, az * az));
              if (d2 < a.rc2_lo) pred = true;
              else if (d2 <= a.rc2_hi) {
                double dx, dy, dz;
                pred = min_image_dist2(pi, pj, R.hb, dx, dy, dz) < a.rc2;
              }
            }
          }
          unsigned long long bits = __ballot(pred) & gmask;
          if (pred) {
            int slot = count + __popcll(bits & lower);
            if (slot < a.nmax) out[slot] = j;
          }
          count += __popcll(bits);
        }
      }
    }
  }
  if (t == 0) {
    a.nnum[row] = count < a.nmax ? count : a.nmax;
    if (count > a.nmax) raise_err(a.err, a.err_code, a.orig[i] + 1, count, a.step);
    else if (a.exact && count != a.nmax) raise_err(a.err, PFM_E_GRAPHENE_NN, a.orig[i] + 1, count, a.step);
  }
}

// slot of the row atom inside each neighbour's own row (Tersoff-Brenner's
// `do l=1,nl%nnum(j); if(nl%nlist(l,j)==i) exit` and graphene's l1 search)
__global__ void k_rev(int nrows, const int* __restrict__ rows, const int* __restrict__ row_of,
                      const int* __restrict__ nnum, const int* __restrict__ nbr, int stride, int* __restrict__ rev,
                      DevErr* err, const int* __restrict__ orig, int step) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int i = rows[r], n = nnum[r];
  for (int p = 0; p < n; ++p) {
    int j = nbr[(size_t)r * stride + p];
    int jr = row_of[j];
    int l = -1;
    if (jr >= 0) {
      int nj = nnum[jr];
      for (int q = 0; q < nj; ++q)
        if (nbr[(size_t)jr * stride + q] == i) { l = q; break; }
    }
    rev[(size_t)r * stride + p] = l;
    if (l < 0) raise_err(err, PFM_E_GRAPHENE_NN, orig[i] + 1, -1, step);
  }
}

// lj1g: every atom's group-local index (the group of the list's second side)
__global__ void k_lj1g_loc(int N, const int* __restrict__ orig, const int* __restrict__ loc, int2* __restrict__ lc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) lc[i].x = loc[orig[i]];
}

// lj1g: group-local index of the first listed neighbour whose atom index is
// larger than the row atom's (the reference's lessnnum boundary, in list
// order = ascending group order); T lanes per row, the minimum over the
// row's entries
template <int T>
__global__ void __launch_bounds__(256) k_lj1g_cut(int nrows, const int* __restrict__ rows,
                                                  const int* __restrict__ nnum, const int* __restrict__ nbr, int stride,
                                                  const int* __restrict__ orig, int2* __restrict__ lc) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int r = gt / T, t = threadIdx.x & (T - 1);
  if (r >= nrows) return;
  int i = rows[r], oi = orig[i], n = nnum[r], c = 0x7fffffff;
  for (int p = t; p < n; p += T) {
    int j = nbr[(size_t)r * stride + p];
    if (orig[j] > oi) { int l = lc[j].x; c = l < c ? l : c; }
  }
#pragma unroll
  for (int o = T / 2; o > 0; o >>= 1) {
    int v = __shfl_xor(c, o, T);
    c = v < c ? v : c;
  }
  if (t == 0) lc[i].y = c;
}

// ljc/morsec metal rows (nl(2)) that have entries, appended in any order
// (each row's force is summed on its own, so the order of the work list does
// not change any result); the count goes after the rows.  One atomic per
// 1024-thread workgroup: wave ballots, then the wave counts' prefix in LDS.
__global__ void __launch_bounds__(1024) k_mark_rows(int nrows, const int* __restrict__ nnum, int* __restrict__ mrows) {
  __shared__ int wc[17];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const bool keep = r < nrows && nnum[r] > 0;
  const unsigned long long bal = __ballot(keep);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) wc[w] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int q = 0; q < nw; ++q) { const int c = wc[q]; wc[q] = tot; tot += c; }
    wc[16] = tot ? atomicAdd(&mrows[nrows], tot) : 0;
  }
  __syncthreads();
  if (keep) mrows[wc[16] + wc[w] + __popcll(bal & ((1ull << lane) - 1ull))] = r;
}

// lj1g: each entry's weight in the gather, [loc(j) < cut(i)] (the pair is in
// i's half list) and [loc(i) < cut(j)] (in j's), packed into bits 30 and 31
// of the entry once per rebuild, so the force pass gathers positions only
__global__ void __launch_bounds__(256) k_lj1g_tag(int nrows, const int* __restrict__ rows,
                                                  const int* __restrict__ nnum, int* __restrict__ nbr, int stride,
                                                  const int2* __restrict__ lc) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int r = gt / 8, t = threadIdx.x & 7;
  if (r >= nrows) return;
  const int2 mi = lc[rows[r]];
  const int n = nnum[r];
  for (int p = t; p < n; p += 8) {
    const int j = nbr[(size_t)r * stride + p];
    const int2 mj = lc[j];
    const unsigned tag = (mj.x < mi.y ? 1u : 0u) | (mi.x < mj.y ? 2u : 0u);
    nbr[(size_t)r * stride + p] = (int)((unsigned)j | (tag << 30));
  }
}

// graphene nearest neighbours from the tb list: entries with |dr| < r_cut_nn,
// exactly three per atom (find_gr_nearest_neighbors, graphenenorm.f90:8-36)
__global__ void k_nn_from_tb(int nrows, const int* __restrict__ rows, const int* __restrict__ tnnum,
                             const int* __restrict__ tnbr, int tstride, const double4* __restrict__ pos,
                             const int* __restrict__ rep, const RepDev* __restrict__ rdev, double rnn,
                             int* __restrict__ nnum, int* __restrict__ nbr, DevErr* err, const int* __restrict__ orig,
                             int step) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int i = rows[r], n = tnnum[r], k = 0;
  const double3 hb = rdev[rep ? rep[i] : 0].hb;
  double4 pi = pos[i];
  for (int p = 0; p < n; ++p) {
    int j = tnbr[(size_t)r * tstride + p];
    double dx, dy, dz;
    double d = sqrt(min_image_dist2(pi, pos[j], hb, dx, dy, dz));
    if (d < rnn) {
      if (k < 3) nbr[(size_t)r * 3 + k] = j;
      ++k;
    }
  }
  nnum[r] = k < 3 ? k : 3;
  if (k != 3) raise_err(err, PFM_E_GRAPHENE_NN, orig[i] + 1, k, step);
}

// Cell-tile list search: one workgroup per cell (grids at least three cells
// wide).  The atoms of the 27 surrounding cells are staged in LDS as FP32
// coordinates unwrapped into the frame of the centre cell's origin, with the
// atom index in w.  Atoms outside group 2, and the padding after the last
// atom, are staged far away, so they never pass the distance test.  With
// CMP (group 2 under half of the atoms: another species' list) only the
// group-2 atoms are staged, compacted in tile order by ballots and a prefix
// over the workgroup's four waves, so the rows walk only real candidates.  Every group-1 atom of the centre cell is a row scanned by T
// lanes; all rows of a wavefront walk the tile in the same order, so their
// LDS reads are broadcasts.  Each lane takes 32 of its candidates (every T-th
// slot) into a register bitmask, two instructions per candidate after the
// distance; the row's lanes then take their output offsets from a prefix sum
// over the T lanes and write their hits.  The FP32 distance decides unless it
// lies within a relative 3e-5 of r_cut**2, where the reference's FP64
// find_distance test (min_image_dist2) on the global positions decides.
// Tiles larger than the LDS buffer are staged in several passes.
// the row of atom i (if it has one) when its tile holds no candidate: what
// the search writes for a row with no entries
__device__ __forceinline__ void empty_row(const SearchArgs& a, const int* __restrict__ row_of, int i) {
  const int row = row_of[i];
  if (row < 0) return;
  a.nnum[row] = 0;
  if (a.nnum2) a.nnum2[row] = 0;
  if ((a.mask[i] & a.g1bit) && a.exact && a.nmax != 0) raise_err(a.err, PFM_E_GRAPHENE_NN, a.orig[i] + 1, 0, a.step);
}

#define TILE_MAX 1536      // staged candidates per pass (24 KB of LDS), a multiple of 32 * T
#define TILE_FAR 1.0e30f   // staged coordinate of the padding

struct TileCells {
  int start[27], cnt[27], off[28];
};

template <int T, bool CMP>
__global__ void __launch_bounds__(256, CMP ? 5 : 6) k_build_tile(SearchArgs a, const int* __restrict__ rank_xyz,
                                                    const int* __restrict__ row_of, int nblocks_real) {
  static_assert(TILE_MAX % (32 * T) == 0, "tile passes hold whole bitmask rounds");
  __shared__ float4 st[TILE_MAX];
  __shared__ TileCells tc;
  int c = xcd_block(nblocks_real);
  if (c < 0) return;
  const int ncen = a.cell_count[c];
  if (ncen == 0 || !(a.cell_grp[c] & a.g1mem)) return;  // no member of group 1 in the cell
  const RepDev& RG = a.rdev[a.cell_rep ? a.cell_rep[c] : 0];
  if (!RG.tile) return;  // searched by k_build
  const int cs = a.cell_start[c];
  const int tid = threadIdx.x;
  // (cells without a row of this list, another species or another
  // replica's interactions, returned above: every group-1 member has a row)
  const int pk = rank_xyz[c];
  const int cc[3] = {pk & 1023, (pk >> 10) & 1023, (pk >> 20) & 1023};
  const double org[3] = {cc[0] * RG.clen.x, cc[1] * RG.clen.y, cc[2] * RG.clen.z};
  const double3 boxL = RG.L, hb = RG.hb;
  bool cand = false;
  if (tid < 27) {
    int o[3] = {tid % 3 - 1, (tid / 3) % 3 - 1, tid / 9 - 1};
    int q[3];
    for (int d = 0; d < 3; ++d) {
      int nd = RG.nc[d];
      q[d] = cc[d] + o[d];
      q[d] = q[d] < 0 ? q[d] + nd : (q[d] >= nd ? q[d] - nd : q[d]);
    }
    int r = a.cell_rank[RG.cell_base + (q[2] * RG.nc[1] + q[1]) * RG.nc[0] + q[0]];
    tc.start[tid] = a.cell_start[r];
    tc.cnt[tid] = a.cell_count[r];
    cand = (a.cell_grp[r] & a.g2bit) != 0;
  }
  if (!__syncthreads_or(cand)) {
    // no group-2 atom in the tile: every row of the cell is empty
    for (int m = tid; m < ncen; m += 256) empty_row(a, row_of, a.bin ? a.bin[cs + m] : cs + m);
    return;
  }
  if (tid == 0) {
    int off = 0;
    for (int k = 0; k < 27; ++k) { tc.off[k] = off; off += tc.cnt[k]; }
    tc.off[27] = off;
  }
  __syncthreads();
  const int R = tc.off[27];
  const float hi = (float)(a.rc2 * (1.0 + 3e-5)), lo = (float)(a.rc2 * (1.0 - 3e-5)), win = hi - lo;

  // stage the group-2 atoms of tile slots [s0, s1) compacted into st (slot
  // m < R belongs to the last tile cell k with off[k] <= m; off is
  // non-decreasing, off[0] = 0), pad them to whole rounds and return that
  // length; every thread calls it.  A thread loads all its slots (every
  // 256th) before the ballots, so the pass waits on one memory latency.
  // dense: every slot was kept, so tile slots and staged slots coincide.
  constexpr int KS = TILE_MAX / 256;
  __shared__ int wsum[KS][4];
  bool dense = true;
  auto stage = [&](int s0, int s1) -> int {
    if constexpr (!CMP) {
      const int np = (s1 - s0 + 32 * T - 1) / (32 * T) * (32 * T);
      for (int m = s0 + tid; m < s0 + np; m += 256) {
        float4 v = make_float4(TILE_FAR, TILE_FAR, TILE_FAR, __int_as_float(-1));
        if (m < s1) {
          int k = 0;
#pragma unroll
          for (int sft = 16; sft > 0; sft >>= 1)
            if (k + sft < 27 && tc.off[k + sft] <= m) k += sft;
          int j = tc.start[k] + (m - tc.off[k]);
          if (a.bin) j = a.bin[j];
          int o0 = k % 3 - 1, o1 = (k / 3) % 3 - 1, o2 = k / 9 - 1;
          double4 p = a.pos[j];
          double x = p.x - org[0], y = p.y - org[1], z = p.z - org[2];
          if (o0 < 0 && x > 0.0) x -= boxL.x; else if (o0 > 0 && x < 0.0) x += boxL.x;
          if (o1 < 0 && y > 0.0) y -= boxL.y; else if (o1 > 0 && y < 0.0) y += boxL.y;
          if (o2 < 0 && z > 0.0) z -= boxL.z; else if (o2 > 0 && z < 0.0) z += boxL.z;
          v.w = __int_as_float(j);
          if (a.mask[j] & a.g2bit) { v.x = (float)x; v.y = (float)y; v.z = (float)z; }
        }
        st[m - s0] = v;
      }
      return np;
    }
    const int lane = tid & 63, w = tid >> 6;
    float4 v[KS];
    bool keep[KS];
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      const int m = s0 + u * 256 + tid;
      keep[u] = false;
      v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (m < s1) {
        int k = 0;
#pragma unroll
        for (int sft = 16; sft > 0; sft >>= 1)
          if (k + sft < 27 && tc.off[k + sft] <= m) k += sft;
        int j = tc.start[k] + (m - tc.off[k]);
        if (a.bin) j = a.bin[j];
        int o0 = k % 3 - 1, o1 = (k / 3) % 3 - 1, o2 = k / 9 - 1;
        double4 p = a.pos[j];
        keep[u] = (a.mask[j] & a.g2bit) != 0;
        // unwrap: the image of atom j next to the centre cell
        double x = p.x - org[0], y = p.y - org[1], z = p.z - org[2];
        if (o0 < 0 && x > 0.0) x -= boxL.x; else if (o0 > 0 && x < 0.0) x += boxL.x;
        if (o1 < 0 && y > 0.0) y -= boxL.y; else if (o1 > 0 && y < 0.0) y += boxL.y;
        if (o2 < 0 && z > 0.0) z -= boxL.z; else if (o2 > 0 && z < 0.0) z += boxL.z;
        v[u] = make_float4((float)x, (float)y, (float)z, __int_as_float(j));
      }
    }
    unsigned long long bal[KS];
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      bal[u] = __ballot(keep[u]);
      if (lane == 0) wsum[u][w] = __popcll(bal[u]);
    }
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      int off = base;
      for (int q = 0; q < w; ++q) off += wsum[u][q];
      if (keep[u]) st[off + __popcll(bal[u] & ((1ull << lane) - 1ull))] = v[u];
      base += wsum[u][0] + wsum[u][1] + wsum[u][2] + wsum[u][3];
    }
    dense = base == s1 - s0;
    const int np = (base + 32 * T - 1) / (32 * T) * (32 * T);
    for (int m = base + tid; m < np; m += 256) st[m] = make_float4(TILE_FAR, TILE_FAR, TILE_FAR, __int_as_float(-1));
    return np;
  };
  const bool single = R <= TILE_MAX;
  int np = 0;
  if (single) np = stage(0, R);

  const int t = tid & (T - 1);
  for (int q0 = 0; q0 < ncen; q0 += 256 / T) {
    const int qq = q0 + tid / T;
    const bool inrow = qq < ncen;
    const int i = a.bin ? a.bin[cs + (inrow ? qq : 0)] : cs + (inrow ? qq : 0);
    const int row = inrow ? row_of[i] : -1;
    const bool active = row >= 0 && (a.mask[i] & a.g1bit);
    const double4 pi = a.pos[i];
    const float px = (float)(pi.x - org[0]), py = (float)(pi.y - org[1]), pz = (float)(pi.z - org[2]);
    int* out = a.nbr + (size_t)(active ? row : 0) * a.stride;
    int count = 0;
    for (int s0 = 0; s0 < R; s0 += TILE_MAX) {
      if (!single) {
        __syncthreads();
        np = stage(s0, R < s0 + TILE_MAX ? R : s0 + TILE_MAX);
      }
      __syncthreads();
      if (!active) continue;
      // the row atom's own staged slot when the pass is dense (else found by index)
      const int own = (!CMP || dense) ? tc.off[13] + qq - s0 : -1;
      for (int c0 = 0; c0 < np; c0 += 32 * T) {
        // bit 31-b <-> candidate slot c0 + t + b*T
        const float4* sv = st + c0 + t;
        // two masks with two instructions each per candidate after the
        // differences: the margin m = hi - d2 as a chain of fused
        // multiply-adds, whose sign bit (set for d2 > hi, and for the far
        // padding's -inf) shifts into the miss mask, and the sign of
        // (hi - lo) - m, set for d2 < lo (clearly inside r_cut)
        unsigned miss = 0u, inside = 0u;
#pragma unroll 8
        for (int b = 0; b < 32; ++b) {
          const float4 v = sv[b * T];
          float dx = v.x - px, dy = v.y - py, dz = v.z - pz;
          const float m = fmaf(-dx, dx, fmaf(-dy, dy, fmaf(-dz, dz, hi)));
          miss = __builtin_amdgcn_alignbit(miss, __float_as_uint(m), 31);
          inside = __builtin_amdgcn_alignbit(inside, __float_as_uint(win - m), 31);
        }
        unsigned hit = ~miss;
        // the row atom is not its own neighbour
        const int ob = own - c0 - t;
        if (ob >= 0 && ob < 32 * T && ob % T == 0) hit &= ~(0x80000000u >> (ob / T));
        // hits inside the FP32 window: the FP64 test decides (with compacted
        // staging that is not dense every hit is also checked for the row atom)
        unsigned chk = CMP && !dense ? hit : hit & ~inside;
        while (chk) {
          const int b = __clz(chk);
          chk &= ~(0x80000000u >> b);
          const float4 v = sv[b * T];
          const int j = __float_as_int(v.w);
          float dx = v.x - px, dy = v.y - py, dz = v.z - pz;
          if (CMP && !dense && j == i) {
            hit &= ~(0x80000000u >> b);
          } else if (fmaf(dx, dx, fmaf(dy, dy, dz * dz)) >= lo) {
            double ex, ey, ez;
            if (!(min_image_dist2(pi, a.pos[j], hb, ex, ey, ez) < a.rc2)) hit &= ~(0x80000000u >> b);
          }
        }
        // offsets of this lane's hits inside the row
        const int nh = __popc(hit);
        int pre = nh;
#pragma unroll
        for (int o = 1; o < T; o <<= 1) {
          int v = __shfl_up(pre, o, T);
          if (t >= o) pre += v;
        }
        const int tot = __shfl(pre, T - 1, T);
        int slot = count + pre - nh;
        while (hit) {
          const int b = __clz(hit);
          hit &= ~(0x80000000u >> b);
          if (slot < a.nmax) out[slot] = __float_as_int(sv[b * T].w);
          ++slot;
        }
        count += tot;
      }
    }
    if (row >= 0 && !active && t == 0) a.nnum[row] = 0;  // an inactive member of a growing group
    if (active && t == 0) {
      a.nnum[row] = count < a.nmax ? count : a.nmax;
      if (count > a.nmax) raise_err(a.err, a.err_code, a.orig[i] + 1, count, a.step);
      else if (a.exact && count != a.nmax) raise_err(a.err, PFM_E_GRAPHENE_NN, a.orig[i] + 1, count, a.step);
    }
  }
}

// Wave-per-cell variant of k_build_tile for sparse grids (a workgroup per
// cell would leave most of its lanes without rows): each of a workgroup's
// four waves searches one cell with its own staged tile of WTILE candidates
// per pass.  The tile order, the tests and the output are those of
// k_build_tile; the tile's cell table lives in lanes 0..26 and is read with
// shuffles, which every lane of the wave takes part in.
#define WTILE 512

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define ENTRY_IDX 0x3fffffff  // staged index without the pair search's group bits
template <int T, bool CMP, bool PAIR, bool QUAD = false>
__global__ void __launch_bounds__(256, 5) k_build_tile_wave(SearchArgs a, const int* __restrict__ rank_xyz,
                                                         const int* __restrict__ row_of, int ncell,
                                                         int nblocks_real) {
  static_assert(WTILE % (32 * T) == 0 && WTILE % 64 == 0, "tile passes hold whole bitmask rounds");
  constexpr int KW = WTILE / 64;
  __shared__ float4 stw[4][WTILE];
  const int lb = xcd_block(nblocks_real);
  if (lb < 0) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = lb * 4 + wv;
  if (c >= ncell) return;
  const int ncen = a.cell_count[c];
  if (ncen == 0 || !(a.cell_grp[c] & a.g1mem)) return;  // no member of group 1 in the cell
  const RepDev& RG = a.rdev[a.cell_rep ? a.cell_rep[c] : 0];
  if (!RG.tile) return;  // searched by k_build
  const int cs = a.cell_start[c];
  float4* st = stw[wv];
  const int pk = rank_xyz[c];
  const int cc[3] = {pk & 1023, (pk >> 10) & 1023, (pk >> 20) & 1023};
  const double org[3] = {cc[0] * RG.clen.x, cc[1] * RG.clen.y, cc[2] * RG.clen.z};
  const double3 boxL = RG.L, hb = RG.hb;
  // the 27 tile cells in lanes 0..26: first atom, count, exclusive offset
  int kst = 0, kcn = 0;
  bool cand = false;
  if (lane < 27) {
    const int o[3] = {lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1};
    int q[3];
    for (int d = 0; d < 3; ++d) {
      int nd = RG.nc[d];
      q[d] = cc[d] + o[d];
      q[d] = q[d] < 0 ? q[d] + nd : (q[d] >= nd ? q[d] - nd : q[d]);
    }
    const int r = a.cell_rank[RG.cell_base + (q[2] * RG.nc[1] + q[1]) * RG.nc[0] + q[0]];
    kst = a.cell_start[r];
    kcn = a.cell_count[r];
    cand = (a.cell_grp[r] & a.g2bit) != 0;
  }
  if (!__any(cand)) {
    for (int m = lane; m < ncen; m += 64) {
      const int i = a.bin ? a.bin[cs + m] : cs + m;
      if (QUAD && (a.mask[i] & a.g1mem_b)) {
        const int row = a.row_of_b[i];
        if (row >= 0) { a.nnum_b1[row] = 0; a.nnum_b2[row] = 0; }
      } else {
        empty_row(a, row_of, i);
      }
    }
    return;
  }
  int inc = kcn;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(inc, o);
    if (lane >= o) inc += v;
  }
  const int koff = inc - kcn;
  const int R = __shfl(inc, 63);
  const int koff13 = __shfl(koff, 13);
  const float hi = (float)(a.rc2 * (1.0 + 3e-5)), lo = (float)(a.rc2 * (1.0 - 3e-5)), win = hi - lo;
  // tile slot m < R: the last cell k with off[k] <= m (every lane takes part)
  auto cell_of = [&](int m) -> int {
    int k = 0;
#pragma unroll
    for (int sft = 16; sft > 0; sft >>= 1) {
      const int kk = k + sft;
      const int o = __shfl(koff, kk < 27 ? kk : 0);
      if (kk < 27 && o <= m) k = kk;
    }
    return k;
  };
  // stage tile slots [s0, s1) (s1 - s0 <= WTILE): every slot, atoms outside
  // group 2 far away; with CMP only group-2 atoms, compacted in tile order.
  // Returns the staged length padded to whole rounds.
  bool dense = true;
  // slots in chunks of KC per lane (KC * 64 per wave): a tile of up to that
  // many candidates waits on one memory latency, and the registers hold one
  // chunk
  constexpr int KC = KW < 4 ? KW : 4;
  auto stage = [&](int s0, int s1) -> int {
    int base = 0;
    for (int cb = 0; cb < KW && s0 + cb * 64 < s1; cb += KC) {
      float4 v[KC];
      bool keep[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int m = s0 + (cb + u) * 64 + lane;
        const int k = cell_of(m < s1 ? m : s0);
        const int ks = __shfl(kst, k), ko = __shfl(koff, k);
        v[u] = make_float4(TILE_FAR, TILE_FAR, TILE_FAR, __int_as_float(-1));
        keep[u] = false;
        if (m < s1) {
          int j = ks + (m - ko);
          if (a.bin) j = a.bin[j];
          const int o0 = k % 3 - 1, o1 = (k / 3) % 3 - 1, o2 = k / 9 - 1;
          const double4 p = a.pos[j];
          keep[u] = (a.mask[j] & a.g2bit) != 0;
          double x = p.x - org[0], y = p.y - org[1], z = p.z - org[2];
          if (o0 < 0 && x > 0.0) x -= boxL.x; else if (o0 > 0 && x < 0.0) x += boxL.x;
          if (o1 < 0 && y > 0.0) y -= boxL.y; else if (o1 > 0 && y < 0.0) y += boxL.y;
          if (o2 < 0 && z > 0.0) z -= boxL.z; else if (o2 > 0 && z < 0.0) z += boxL.z;
          if constexpr (PAIR) {
            // bit 30: in list 1's candidate group, bit 31: in list 2's
            const uint32_t mj = a.mask[j];
            v[u].w = __uint_as_float((uint32_t)j | ((mj & a.g2bit1) ? 0x40000000u : 0u) |
                                     ((mj & a.g2bit2) ? 0x80000000u : 0u));
          } else {
            v[u].w = __int_as_float(j);
          }
          if (keep[u]) { v[u].x = (float)x; v[u].y = (float)y; v[u].z = (float)z; }
        }
      }
      if constexpr (!CMP) {
#pragma unroll
        for (int u = 0; u < KC; ++u) {
          const int m = s0 + (cb + u) * 64 + lane;
          if (m < s1) st[m - s0] = v[u];
        }
      } else {
#pragma unroll
        for (int u = 0; u < KC; ++u) {
          const unsigned long long bal = __ballot(keep[u]);
          if (keep[u]) st[base + __popcll(bal & ((1ull << lane) - 1ull))] = v[u];
          base += __popcll(bal);
        }
      }
    }
    if constexpr (!CMP) base = s1 - s0;
    else dense = base == s1 - s0;
    const int np = (base + 32 * T - 1) / (32 * T) * (32 * T);
    for (int m = base + lane; m < np; m += 64) st[m] = make_float4(TILE_FAR, TILE_FAR, TILE_FAR, __int_as_float(-1));
    return np;
  };
  const bool single = R <= WTILE;
  int np = 0;
  if (single) np = stage(0, R);
  wave_sync();

  const int t = lane & (T - 1);
  for (int q0 = 0; q0 < ncen; q0 += 64 / T) {
    const int qq = q0 + lane / T;
    const bool inrow = qq < ncen;
    const int i = a.bin ? a.bin[cs + (inrow ? qq : 0)] : cs + (inrow ? qq : 0);
    const uint32_t mi = a.mask[i];
    // QUAD: rows of side b take its row index and its two lists
    const bool sb = QUAD && (mi & a.g1mem_b) != 0;
    const int row = inrow ? (sb ? a.row_of_b[i] : row_of[i]) : -1;
    const bool active = row >= 0 && (mi & (sb ? a.g1bit_b : a.g1bit));
    const double4 pi = a.pos[i];
    const float px = (float)(pi.x - org[0]), py = (float)(pi.y - org[1]), pz = (float)(pi.z - org[2]);
    int* const nbr1 = sb ? a.nbr_b1 : a.nbr;
    int* const nbr2 = sb ? a.nbr_b2 : a.nbr2;
    int* const nnum1 = sb ? a.nnum_b1 : a.nnum;
    int* const nnum2 = sb ? a.nnum_b2 : a.nnum2;
    const int nmax1 = sb ? a.nmax_b1 : a.nmax, nmax2 = sb ? a.nmax_b2 : a.nmax2;
    int* out = nbr1 + (size_t)(active ? row : 0) * (sb ? a.stride_b1 : a.stride);
    int* out2 = PAIR ? nbr2 + (size_t)(active ? row : 0) * (sb ? a.stride_b2 : a.stride2) : nullptr;
    int count = 0, count2 = 0;
    // the row's hits of one round into one list: offsets from a prefix sum
    // over the row's lanes, entries without the group bits
    auto put = [&](unsigned hit, const float4* sv, int* o, int nmax, int& cnt) {
      const int nh = __popc(hit);
      int pre = nh;
#pragma unroll
      for (int d = 1; d < T; d <<= 1) {
        int w = __shfl_up(pre, d, T);
        if (t >= d) pre += w;
      }
      const int tot = __shfl(pre, T - 1, T);
      int slot = cnt + pre - nh;
      while (hit) {
        const int b = __clz(hit);
        hit &= ~(0x80000000u >> b);
        if (slot < nmax) o[slot] = __float_as_int(sv[b * T].w) & ENTRY_IDX;
        ++slot;
      }
      cnt += tot;
    };
    for (int s0 = 0; s0 < R; s0 += WTILE) {
      if (!single) {
        wave_sync();
        np = stage(s0, R < s0 + WTILE ? R : s0 + WTILE);
        wave_sync();
      }
      if (!active) continue;
      // the row atom's own staged slot when the pass is dense (else found by index)
      const int own = (!CMP || dense) ? koff13 + qq - s0 : -1;
      for (int c0 = 0; c0 < np; c0 += 32 * T) {
        const float4* sv = st + c0 + t;
        unsigned miss = 0u, inside = 0u;  // as in k_build_tile
        unsigned m1 = 0u, m2 = 0u;        // PAIR: candidates in list 1's / list 2's group
#pragma unroll 4
        for (int b = 0; b < 32; ++b) {
          const float4 w = sv[b * T];
          float dx = w.x - px, dy = w.y - py, dz = w.z - pz;
          const float m = fmaf(-dx, dx, fmaf(-dy, dy, fmaf(-dz, dz, hi)));
          miss = __builtin_amdgcn_alignbit(miss, __float_as_uint(m), 31);
          inside = __builtin_amdgcn_alignbit(inside, __float_as_uint(win - m), 31);
          if constexpr (PAIR) {
            const uint32_t wb = __float_as_uint(w.w);
            m1 = __builtin_amdgcn_alignbit(m1, wb << 1, 31);
            m2 = __builtin_amdgcn_alignbit(m2, wb, 31);
          }
        }
        unsigned hit = ~miss;
        const int ob = own - c0 - t;
        if (ob >= 0 && ob < 32 * T && ob % T == 0) hit &= ~(0x80000000u >> (ob / T));
        unsigned chk = CMP && !dense ? hit : hit & ~inside;
        while (chk) {
          const int b = __clz(chk);
          chk &= ~(0x80000000u >> b);
          const float4 w = sv[b * T];
          const int j = __float_as_int(w.w) & ENTRY_IDX;
          float dx = w.x - px, dy = w.y - py, dz = w.z - pz;
          if (CMP && !dense && j == i) {
            hit &= ~(0x80000000u >> b);
          } else if (fmaf(dx, dx, fmaf(dy, dy, dz * dz)) >= lo) {
            double ex, ey, ez;
            if (!(min_image_dist2(pi, a.pos[j], hb, ex, ey, ez) < a.rc2)) hit &= ~(0x80000000u >> b);
          }
        }
        if constexpr (PAIR) {
          put(hit & m1, sv, out, nmax1, count);
          put(hit & m2, sv, out2, nmax2, count2);
        } else {
          put(hit, sv, out, a.nmax, count);
        }
      }
    }
    if (row >= 0 && !active && t == 0) {
      nnum1[row] = 0;
      if (PAIR) nnum2[row] = 0;
    }
    if (active && t == 0) {
      nnum1[row] = count < nmax1 ? count : nmax1;
      if (count > nmax1) raise_err(a.err, a.err_code, a.orig[i] + 1, count, a.step);
      else if (a.exact && count != nmax1) raise_err(a.err, PFM_E_GRAPHENE_NN, a.orig[i] + 1, count, a.step);
      if (PAIR) {
        nnum2[row] = count2 < nmax2 ? count2 : nmax2;
        if (count2 > nmax2) raise_err(a.err, a.err_code, a.orig[i] + 1, count2, a.step);
      }
    }
  }
}

template <int T>
static void launch_build_T(pfm_ctx* h, SearchArgs& a, const int* row_of) {
  long threads = (long)a.nrows * T;
  int nb = (int)((threads + 255) / 256);
  PFM_LAUNCH(h, "build_rows", k_build<T>, pad8(nb), 256, a, nb);
}

static void launch_build(pfm_ctx* h, SearchArgs& a, const int* row_of, bool sparse, bool sparse_cells) {
  if (h->any_tile && sparse_cells) {
    const int nc = h->ncell, nb = (nc + 3) / 4;
    if (sparse) {
      if (a.nmax <= 8) PFM_LAUNCH(h, "build_list", (k_build_tile_wave<4, true, false>), pad8(nb), 256, a, h->cell_xyz, row_of, nc, nb);
      else PFM_LAUNCH(h, "build_list", (k_build_tile_wave<8, true, false>), pad8(nb), 256, a, h->cell_xyz, row_of, nc, nb);
    } else {
      if (a.nmax <= 8) PFM_LAUNCH(h, "build_list", (k_build_tile_wave<4, false, false>), pad8(nb), 256, a, h->cell_xyz, row_of, nc, nb);
      else PFM_LAUNCH(h, "build_list", (k_build_tile_wave<8, false, false>), pad8(nb), 256, a, h->cell_xyz, row_of, nc, nb);
    }
  } else if (h->any_tile) {
    int nb = h->ncell;
    if (sparse) {
      if (a.nmax <= 8) PFM_LAUNCH(h, "build_list", (k_build_tile<4, true>), pad8(nb), 256, a, h->cell_xyz, row_of, nb);
      else PFM_LAUNCH(h, "build_list", (k_build_tile<8, true>), pad8(nb), 256, a, h->cell_xyz, row_of, nb);
    } else {
      if (a.nmax <= 8) PFM_LAUNCH(h, "build_list", (k_build_tile<4, false>), pad8(nb), 256, a, h->cell_xyz, row_of, nb);
      else PFM_LAUNCH(h, "build_list", (k_build_tile<8, false>), pad8(nb), 256, a, h->cell_xyz, row_of, nb);
    }
  }
  if (h->any_rows) {
    if (a.nmax <= 8) launch_build_T<4>(h, a, row_of);
    else launch_build_T<8>(h, a, row_of);
  }
}

// Build list k (0-based) of interaction ii.  For lj/ljc/morsec, list 1 is the
// converse of list 0: same r_cut and schedule (md_interactions.f90:152-155),
// found by searching from the group-2 side, which gives the same set because
// find_distance's dr2 is symmetric in its two atoms.
// the search arguments of list k of interaction ii
static SearchArgs search_args(pfm_ctx* h, int ii, int k, int step) {
  Inter& it = h->inter[ii];
  ListDev& L = it.nl[k];
  SearchArgs a{};
  a.nrows = L.nrows;
  a.rows = h->grp.rows[L.g1];
  a.g1bit = 1u << (L.g1 - 1);
  a.g1mem = 1u << ((h->grp.shadow[L.g1] ? h->grp.shadow[L.g1] : L.g1) - 1);
  a.g2bit = 1u << (L.g2 - 1);
  a.pos = h->pos4;
  a.mask = h->mask;
  a.orig = h->orig;
  a.rdev = h->rdev;
  a.rep = h->rep;
  a.cell_rep = h->cell_rep;
  a.bin = h->bin;
  double rc = (k == 1 && has_converse(it.kind)) ? it.nl[0].rcut : L.rcut;
  a.rc2 = rc * rc;
  a.rc2_lo = a.rc2 * (1.0 - 1e-12);
  a.rc2_hi = a.rc2 * (1.0 + 1e-12);
  a.cell_rank = h->cell_rank;
  a.cell_start = h->cell_start;
  a.cell_count = h->cell_count;
  a.cell_grp = h->cell_grp;
  a.nmax = L.nmax;
  a.stride = L.stride;
  a.nnum = L.nnum;
  a.nbr = L.nbr;
  a.err = h->derr;
  a.step = step - h->step_base;
  a.err_code = PFM_E_OVERFLOW;
  a.exact = (k == 2) ? 1 : 0;
  return a;
}

// what follows a list's search: tb reverse slots, the ljc/morsec metal work
// list, lj1g half-list tags
static void list_post(pfm_ctx* h, int ii, int k, int step) {
  Inter& it = h->inter[ii];
  ListDev& L = it.nl[k];
  int nb = (L.nrows + 255) / 256;
  if (it.kind == PFM_TB || it.kind == PFM_REBOSC || (k == 2 && !(int)it.p[it.kind == PFM_LJC ? 5 : 6])) {
    PFM_LAUNCH(h, "list_rev", k_rev, nb, 256, L.nrows, h->grp.rows[L.g1], h->grp.row_of[L.g1], L.nnum, L.nbr,
               L.stride, L.rev, h->derr, h->orig, step - h->step_base);
  }
  if ((it.kind == PFM_LJC || it.kind == PFM_MORSEC) && k == 1) {
    PFM_LAUNCH(h, "list_rows", k_fill_int, 1, 64, 1, it.mrows + L.nrows, 0);
    PFM_LAUNCH(h, "list_rows", k_mark_rows, (L.nrows + 1023) / 1024, 1024, L.nrows, L.nnum, it.mrows);
  }
  if (it.kind == PFM_LJ1G) {
    PFM_LAUNCH(h, "list_cut", k_lj1g_loc, (h->N + 255) / 256, 256, h->N, h->orig, h->grp.loc[L.g2], it.lc);
    PFM_LAUNCH(h, "list_cut", k_lj1g_cut<8>, (int)(((long)L.nrows * 8 + 255) / 256), 256, L.nrows, h->grp.rows[L.g1],
               L.nnum, L.nbr, L.stride, h->orig, it.lc);
    PFM_LAUNCH(h, "list_cut", k_lj1g_tag, (int)(((long)L.nrows * 8 + 255) / 256), 256, L.nrows, h->grp.rows[L.g1],
               L.nnum, L.nbr, L.stride, it.lc);
  }
  L.built = true;
}

int launch_build_list(pfm_ctx* h, int ii, int k, int step) {
  ListDev& L = h->inter[ii].nl[k];
  if (L.nrows == 0) { L.built = true; return PFM_OK; }
  SearchArgs a = search_args(h, ii, k, step);
  // candidates of another species: stage only those
  launch_build(h, a, h->grp.row_of[L.g1], 2L * h->grp.size[L.g2] < (long)h->N, L.sparse_cells);
  list_post(h, ii, k, step);
  return PFM_OK;
}

// Two lists of one pair (ListDev::pair): one wavefront-per-cell pass stages
// the union of their candidate groups once and writes both (replicas that
// take the row search run it per list).
int launch_build_pair(pfm_ctx* h, int ii, int k, int jj, int kk, int step) {
  ListDev& L1 = h->inter[ii].nl[k];
  ListDev& L2 = h->inter[jj].nl[kk];
  SearchArgs a = search_args(h, ii, k, step);
  SearchArgs b = search_args(h, jj, kk, step);
  const int* row_of = h->grp.row_of[L1.g1];
  if (h->any_tile && L1.nrows > 0) {
    SearchArgs p = a;
    p.g2bit1 = a.g2bit;
    p.g2bit2 = b.g2bit;
    p.g2bit = a.g2bit | b.g2bit;
    p.nmax2 = b.nmax;
    p.stride2 = b.stride;
    p.nnum2 = b.nnum;
    p.nbr2 = b.nbr;
    const int nc = h->ncell, nb = (nc + 3) / 4;
    if (L1.pair_sparse)
      PFM_LAUNCH(h, "build_list", (k_build_tile_wave<8, true, true>), pad8(nb), 256, p, h->cell_xyz, row_of, nc, nb);
    else
      PFM_LAUNCH(h, "build_list", (k_build_tile_wave<8, false, true>), pad8(nb), 256, p, h->cell_xyz, row_of, nc, nb);
  }
  if (h->any_rows) {
    if (L1.nrows > 0) launch_build_T<8>(h, a, row_of);
    if (L2.nrows > 0) launch_build_T<8>(h, b, row_of);
  }
  list_post(h, ii, k, step);
  list_post(h, jj, kk, step);
  return PFM_OK;
}

// Two pairs in one pass (ListDev::quad_*): the lead pair (ii, k) + (jj, kk)
// and a second pair with rows of another group whose lists take the same two
// candidate groups, (b1 for the lead's group, b2 for its partner's).
int launch_build_quad(pfm_ctx* h, int ii, int k, int jj, int kk, int b1i, int b1k, int b2i, int b2k, int step) {
  ListDev& L1 = h->inter[ii].nl[k];
  ListDev& B1 = h->inter[b1i].nl[b1k];
  SearchArgs a = search_args(h, ii, k, step);
  SearchArgs b = search_args(h, jj, kk, step);
  SearchArgs c = search_args(h, b1i, b1k, step);
  SearchArgs d = search_args(h, b2i, b2k, step);
  const int* row_of = h->grp.row_of[L1.g1];
  const int* row_of_b = h->grp.row_of[B1.g1];
  if (h->any_tile && (L1.nrows > 0 || B1.nrows > 0)) {
    SearchArgs p = a;
    p.g2bit1 = a.g2bit;
    p.g2bit2 = b.g2bit;
    p.g2bit = a.g2bit | b.g2bit;
    p.nmax2 = b.nmax; p.stride2 = b.stride; p.nnum2 = b.nnum; p.nbr2 = b.nbr;
    p.g1bit_b = c.g1bit;
    p.g1mem_b = c.g1mem;
    p.g1mem = a.g1mem | c.g1mem;
    p.row_of_b = row_of_b;
    p.nmax_b1 = c.nmax; p.stride_b1 = c.stride; p.nnum_b1 = c.nnum; p.nbr_b1 = c.nbr;
    p.nmax_b2 = d.nmax; p.stride_b2 = d.stride; p.nnum_b2 = d.nnum; p.nbr_b2 = d.nbr;
    const int nc = h->ncell, nb = (nc + 3) / 4;
    if (L1.pair_sparse)
      PFM_LAUNCH(h, "build_list", (k_build_tile_wave<8, true, true, true>), pad8(nb), 256, p, h->cell_xyz, row_of, nc, nb);
    else
      PFM_LAUNCH(h, "build_list", (k_build_tile_wave<8, false, true, true>), pad8(nb), 256, p, h->cell_xyz, row_of, nc, nb);
  }
  if (h->any_rows) {
    if (L1.nrows > 0) { launch_build_T<8>(h, a, row_of); launch_build_T<8>(h, b, row_of); }
    if (B1.nrows > 0) { launch_build_T<8>(h, c, row_of_b); launch_build_T<8>(h, d, row_of_b); }
  }
  list_post(h, ii, k, step);
  list_post(h, jj, kk, step);
  list_post(h, b1i, b1k, step);
  list_post(h, b2i, b2k, step);
  return PFM_OK;
}

// nl(3) of ljc/morsec copied from the tb interaction's freshly rebuilt list
int launch_nn_from_tb(pfm_ctx* h, int ii, int step) {
  Inter& it = h->inter[ii];
  Inter& tb = h->inter[it.nn_from_tb];
  ListDev& L = it.nl[2];
  ListDev& T = tb.nl[0];
  if (L.nrows == 0) return PFM_OK;
  int nb = (L.nrows + 255) / 256;
  PFM_LAUNCH(h, "nn_from_tb", k_nn_from_tb, nb, 256, L.nrows, h->grp.rows[L.g1], T.nnum, T.nbr, T.stride, h->pos4,
             h->rep, h->rdev, L.rcut, L.nnum, L.nbr, h->derr, h->orig, step - h->step_base);
  bool simplified = (int)it.p[it.kind == PFM_LJC ? 5 : 6];
  if (!simplified)
    PFM_LAUNCH(h, "list_rev", k_rev, nb, 256, L.nrows, h->grp.rows[L.g1], h->grp.row_of[L.g1], L.nnum, L.nbr,
               L.stride, L.rev, h->derr, h->orig, step - h->step_base);
  L.built = true;
  return PFM_OK;
}
