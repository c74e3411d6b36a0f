/*
 * pfm_oracle.c -- CPU FP64 restatement of the PFMDS per-step MD path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle of the HIP path, and the CPU
 * baseline timed by bench.py).  Nothing in the product loads this file.
 *
 * It follows the reference's own algorithm and loop order, file by file:
 *   neighbour lists   code_source/MOLECULAR_DYNAMICS/md_neighbours.f90
 *   graphene normals  code_source/INTERACTION_POTENTIALS/graphenenorm.f90
 *   cutoffs           code_source/INTERACTION_POTENTIALS/cut_off_function.f90, cut_off_poly.f90
 *   potentials        code_source/INTERACTION_POTENTIALS/{LennardJones,LennardJones_1g,
 *                     LennardJonesCosine,MorseCosine,RosatoGuillopeLegrand,TersoffBrenner}.f90
 *   dispatch          code_source/MOLECULAR_DYNAMICS/md_interactions.f90:138-271
 *   integrators, NHC  code_source/MOLECULAR_DYNAMICS/md_integrators.f90
 *   reductions        code_source/MOLECULAR_DYNAMICS/md_general.f90
 *   step order        code_source/MOLECULAR_DYNAMICS/md_simulation.f90:114-231
 *   group growth      change_particle_group_N, md_general.f90:82-94
 * It keeps the reference's data layout (padded per-atom lists storing dr and
 * |dr|, brute-force O(N1*N2) search, serial converse transpose), so it is
 * also the "reference algorithm" CPU baseline.
 *
 * Differences from the reference, all at the rounding level:
 *   - energy sums are formed per row and then added in row order (the
 *     reference adds per-thread partials under OMP ATOMIC, in thread order);
 *   - lj1g forces are accumulated serially (the reference merges per-thread
 *     private arrays under OMP ATOMIC, LennardJones_1g.f90:63-99);
 *   - the Tersoff-Brenner 2.d0 literals (TersoffBrenner.f90:101) are FP64 here;
 *     under gfortran -fdefault-real-8 they are REAL(16).
 * Compile with -ffp-contract=off so a*b+c is never fused, as in the
 * reference's default gfortran build.
 *
 * Parity pins (tests/test_oracle_*.py): brute-force neighbour sets (numpy),
 * closed-form energies (dimers, perfect graphene, fcc shell sums),
 * forces = -grad E by torch float64 autograd of independently written
 * energies, the lj1g cutoff quirk as a linear combination of two autograd
 * terms, NHC conserved quantity, NVE reversibility and energy conservation,
 * nvms convergence to the analytic LJ dimer minimum.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "pfm_oracle.h"
#include "../include/pfm.h"

#define MAXG 32
#define MAXI 16
#define MAXNHC 8
/* pi as the cutoff module writes it (cut_off_function.f90:8,20) */
#define OPI 3.14159265358979
/* unit constants, evaluated left to right as md_general.f90:165,304 */
#define MASS_COEF (1.6605389217 / 1.6021765654 * 100.0)
#define KT_A_DEGREE (1.3806488 / 1.6021765654 * (1.0 / 10000.0))

typedef struct { int n, size; int* idx; } ogroup; /* 0-based atom indices; n active of size */

typedef struct { int from, to, ts1, ts2, frec; } ochange;
#define MAXCHANGE 8

typedef struct {
  int N, nmax, period;
  double rcut;
  int *nlist, *nnum, *lessnnum, *pidx; /* nlist: 0-based group2-local */
  double *dr, *moddr;
} onl;

typedef struct {
  int kind, nl_n, g[6];
  double p[24];
  onl nl[3];
  double energy;
  double* gr_norm; /* 3*nl[2].N, graphene normals for ljc/morsec */
  /* derived parameters */
  double c6, c12, c6t6, c12t12, c02, d02;
} ointer;

typedef struct {
  int M, L, group;
  double x[16], v[16], q[16];
  double temperature, s, e;
} onhc;

struct octx {
  int N;
  double box[3], hbox[3];
  double *pos, *vel, *mass, *force; /* 3N, N */
  ogroup groups[MAXG + 1];
  int ninter;
  ointer inter[MAXI];
  int nnhc;
  onhc nhc[MAXNHC];
  int g_all_moving, g_xyz, g_z, g_all, zm_period, invert_z;
  int nchange;
  ochange change[MAXCHANGE];
  double pe, pe_prev, ke, temp, sim_time;
  double ts[4];
  char err[512];
};

static int oerr(octx* h, int code, const char* msg) {
  snprintf(h->err, sizeof h->err, "%s", msg);
  return code;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* find_distance (md_general.f90:423-441): dr = vec2-vec1, minimum image by
 * the sign formula, dr2 = dx*dx+dy*dy+dz*dz. */
static inline double fdist(const double* a, const double* b, const double* hb, double dr[3]) {
  for (int k = 0; k < 3; ++k) {
    double d = b[k] - a[k];
    d = d - hb[k] * (copysign(1.0, d - hb[k]) + copysign(1.0, d + hb[k]));
    dr[k] = d;
  }
  return dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
}

/* cosine cutoff f_cut / df_cut (cut_off_function.f90:6-28); df_cut is f'/r */
static inline double fcut(double r, double R1, double R2) {
  if (r < R1) return 1.0;
  if (r < R2) return (1.0 + cos(OPI * (r - R1) / (R2 - R1))) / 2;
  return 0.0;
}
static inline double dfcut(double r, double R1, double R2) {
  if (r < R1) return 0.0;
  if (r < R2) return -sin(OPI * (r - R1) / (R2 - R1)) * OPI / (R2 - R1) / r / 2;
  return 0.0;
}
/* polynomial cutoff f_dfr_cut (cut_off_poly.f90:30-43), including the
 * missing 1/(R2-R1) factor of dfr */
static inline void fdfr_poly(double r, double R1, double R2, double* f, double* dfr) {
  double t = r;
  t = t > R1 ? t : R1;
  t = t < R2 ? t : R2;
  double x = (t - R1) / (R2 - R1), x2 = x * x;
  *f = 1.0 + x2 * x * (-10.0 + 15.0 * x - 6.0 * x2);
  *dfr = t * x2 * (-30.0 + 60.0 * x - 30.0 * x2);
}
/* polynomial f_cut (cut_off_poly.f90:6-17), used by LJ1g_energy */
static inline double fcut_poly(double r, double R1, double R2) {
  if (r > R1 && r < R2) {
    double a = r - R1, w = R2 - R1;
    return 1.0 + (-10.0 * a * a * a * w * w + 15.0 * a * a * a * a * w - 6.0 * a * a * a * a * a) /
                     (w * w * w * w * w);
  } else if (r <= R1) {
    return 1.0;
  }
  return 0.0;
}
static inline double pow6(double x) { double x2 = x * x; return x2 * x2 * x2; }

/* ------------------------------------------------------------------ setup */

int orc_create(octx** out, int n, const double box[3], const double* pos, const double* vel,
               const double* mass) {
  octx* h = (octx*)calloc(1, sizeof(octx));
  if (!h) return PFM_E_INVALID;
  *out = h;
  if (n <= 0) return oerr(h, PFM_E_INVALID, "error: particles_num must be positive");
  h->N = n;
  for (int k = 0; k < 3; ++k) { h->box[k] = box[k]; h->hbox[k] = 0.5 * box[k]; }
  h->pos = (double*)malloc(sizeof(double) * 3 * n);
  h->vel = (double*)malloc(sizeof(double) * 3 * n);
  h->force = (double*)calloc(3 * (size_t)n, sizeof(double));
  h->mass = (double*)malloc(sizeof(double) * n);
  memcpy(h->pos, pos, sizeof(double) * 3 * n);
  memcpy(h->vel, vel, sizeof(double) * 3 * n);
  memcpy(h->mass, mass, sizeof(double) * n);
  h->zm_period = 1;
  return PFM_OK;
}

int orc_set_group(octx* h, int gid, int n, const int* idx) {
  if (gid < 1 || gid > MAXG) return oerr(h, PFM_E_INVALID, "error: group number out of range");
  ogroup* g = &h->groups[gid];
  free(g->idx);
  g->n = n;
  g->size = n;
  g->idx = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    if (idx[i] < 1 || idx[i] > h->N) return oerr(h, PFM_E_INVALID, "error: group index out of range");
    g->idx[i] = idx[i] - 1;
  }
  return PFM_OK;
}

int orc_set_roles(octx* h, int all_moving, int xyz_moving, int z_moving, int all_atoms, int zmp,
                  int invert_z) {
  if (zmp < 1) return oerr(h, PFM_E_INVALID, "error: zero_momentum_period must be >= 1");
  h->g_all_moving = all_moving; h->g_xyz = xyz_moving; h->g_z = z_moving; h->g_all = all_atoms;
  h->zm_period = zmp; h->invert_z = invert_z;
  return PFM_OK;
}

/* create_neighbour_list (md_neighbours.f90:9-28) */
static void nl_alloc(onl* nl) {
  int N = nl->N > 0 ? nl->N : 1, m = nl->nmax > 0 ? nl->nmax : 1;
  nl->pidx = (int*)calloc(N, sizeof(int));
  nl->nnum = (int*)calloc(N, sizeof(int));
  nl->lessnnum = (int*)calloc(N, sizeof(int));
  nl->nlist = (int*)calloc((size_t)N * m, sizeof(int));
  nl->dr = (double*)calloc((size_t)N * m * 3, sizeof(double));
  nl->moddr = (double*)calloc((size_t)N * m, sizeof(double));
}

int orc_add_interaction(octx* h, int kind, const double* prm, int np, int nl_n, const int* gn,
                        const int* nmax, const double* rcut, const int* period) {
  static const int need_np[9] = {0, 4, 4, 6, 7, 7, 10, 0, 18};
  static const int need_nl[9] = {0, 2, 1, 3, 3, 1, 1, 0, 1};
  if (kind < 1 || kind > 8 || kind == PFM_SUR) return oerr(h, PFM_E_UNKNOWN_KIND, "error: unknown interaction name");
  if (h->ninter >= MAXI) return oerr(h, PFM_E_INVALID, "error: too many interactions");
  if (np < need_np[kind] || nl_n != need_nl[kind])
    return oerr(h, PFM_E_INVALID, "error: wrong parameter or list count for interaction");
  ointer* it = &h->inter[h->ninter];
  memset(it, 0, sizeof *it);
  it->kind = kind; it->nl_n = nl_n;
  for (int i = 0; i < np && i < 24; ++i) it->p[i] = prm[i];
  for (int j = 0; j < nl_n; ++j) {
    it->g[2 * j] = gn[2 * j]; it->g[2 * j + 1] = gn[2 * j + 1];
    if (gn[2 * j] < 1 || gn[2 * j] > MAXG || gn[2 * j + 1] < 1 || gn[2 * j + 1] > MAXG)
      return oerr(h, PFM_E_INVALID, "error: interaction group out of range");
    if (period[j] < 1) return oerr(h, PFM_E_INVALID, "error: update_period must be >= 1");
    it->nl[j].nmax = nmax[j]; it->nl[j].rcut = rcut[j]; it->nl[j].period = period[j];
    it->nl[j].N = h->groups[gn[2 * j]].n; /* md_interactions.f90:130 */
    nl_alloc(&it->nl[j]);
  }
  if (kind == PFM_LJ1G) { /* read_LJ1g_parameters (LennardJones_1g.f90:20-23) */
    double eps = it->p[0], sig = it->p[1];
    it->c6 = 4.0 * eps * pow6(sig);
    it->c12 = 4.0 * eps * pow6(sig) * pow6(sig);
    it->c6t6 = 6.0 * 4.0 * eps * pow6(sig);
    it->c12t12 = 12.0 * 4.0 * eps * pow6(sig) * pow6(sig);
  }
  if (kind == PFM_TB) { it->c02 = it->p[6] * it->p[6]; it->d02 = it->p[7] * it->p[7]; }
  if (kind == PFM_LJC || kind == PFM_MORSEC) /* allocate_graphene_norm (md_interactions.f90:180-193) */
    it->gr_norm = (double*)calloc(3 * (size_t)(it->nl[2].N > 0 ? it->nl[2].N : 1), sizeof(double));
  h->ninter++;
  return PFM_OK;
}

/* set_nose_hoover_chain (md_integrators.f90:182-198) */
int orc_add_nhc(octx* h, int group, double T, int M, double q1) {
  if (h->nnhc >= MAXNHC || M < 1 || M > 16 || group < 1 || group > MAXG)
    return oerr(h, PFM_E_INVALID, "error: bad nhc definition");
  int L = h->groups[group].n;
  if (L < 1 || q1 < 0.0 || T < 0.0) return oerr(h, PFM_E_NHC_PARAMS, "error: wrong nhc parameters");
  onhc* c = &h->nhc[h->nnhc++];
  memset(c, 0, sizeof *c);
  c->M = M; c->L = L; c->group = group; c->temperature = T; c->s = 1.0;
  c->q[0] = q1;
  for (int i = 1; i < M; ++i) c->q[i] = c->q[0] / (3.0 * c->L);
  return PFM_OK;
}

/* Positions x[M] and velocities v[M] of chain c (1-based): the
 * nose_hoover_chain state (md_general.f90:24-28); mirrors pfm_get_chain. */
int orc_get_chain(octx* h, int c, double* x, double* v) {
  if (c < 1 || c > h->nnhc) return oerr(h, PFM_E_INVALID, "error: no such chain");
  const onhc* d = &h->nhc[c - 1];
  for (int i = 0; i < d->M; ++i) { x[i] = d->x[i]; v[i] = d->v[i]; }
  return PFM_OK;
}

/* One runtime group change of the settings (md_simulation.f90:63-70). */
int orc_add_group_change(octx* h, int from, int to, int ts1, int ts2, int frec) {
  if (h->nchange >= MAXCHANGE || from < 1 || from > MAXG || to < 1 || to > MAXG || frec < 1)
    return oerr(h, PFM_E_INVALID, "error: bad group change");
  ochange* c = &h->change[h->nchange++];
  c->from = from; c->to = to; c->ts1 = ts1; c->ts2 = ts2; c->frec = frec;
  return PFM_OK;
}

int orc_pe_change(octx* h, double* de) {
  *de = h->pe - h->pe_prev;
  return PFM_OK;
}

int orc_group_count(octx* h, int gid, int* n) {
  if (gid < 0 || gid > MAXG) return oerr(h, PFM_E_INVALID, "error: group number out of range");
  *n = h->groups[gid].n;
  return PFM_OK;
}

/* change_particle_group_N (md_general.f90:82-94): up to step ts1 the group
 * holds as many of its atoms as `from` does, one more at ts1, then one more
 * every frec steps before ts2, never more than it has */
static void change_group_n(ogroup* g, int step, const ochange* c, const ogroup* from) {
  if (step <= c->ts1) {
    g->n = from->n;
    if (step == c->ts1) g->n = g->n + 1;
  } else if (step < c->ts2 && (step - c->ts1) % c->frec == 0) {
    g->n = g->n + 1;
  }
  if (g->n > g->size) g->n = g->size;
}

/* --------------------------------------------------------- neighbour lists */

/* find_neighbours (md_neighbours.f90:56-100) */
static int find_neighbours(octx* h, onl* nl, const ogroup* g1, const ogroup* g2) {
  if (g1->n > nl->N) return oerr(h, PFM_E_INVALID, "error: group1%N>nl%N");
  int bad = 0, bad_ind = 0, bad_n = 0;
  double rc2 = nl->rcut * nl->rcut;
#pragma omp parallel for schedule(static)
  for (int ind = 0; ind < g1->n; ++ind) {
    int i = g1->idx[ind];
    nl->pidx[ind] = i;
    int nn = 0, less = -1;
    for (int jnd = 0; jnd < g2->n; ++jnd) {
      int j = g2->idx[jnd];
      if (i == j) continue;
      double dr[3];
      double dr2 = fdist(&h->pos[3 * i], &h->pos[3 * j], h->hbox, dr);
      if (dr2 < rc2) {
        if (less == -1 && i < j) less = nn;
        nn++;
        if (nn > nl->nmax) {
#pragma omp critical
          { if (!bad) { bad = 1; bad_ind = ind + 1; bad_n = nn; } }
          break;
        }
        size_t e = (size_t)ind * nl->nmax + (nn - 1);
        nl->nlist[e] = jnd;
        nl->dr[3 * e] = dr[0]; nl->dr[3 * e + 1] = dr[1]; nl->dr[3 * e + 2] = dr[2];
        nl->moddr[e] = sqrt(dr2);
      }
    }
    nl->lessnnum[ind] = (less == -1) ? nn : less;
    nl->nnum[ind] = nn > nl->nmax ? nl->nmax : nn;
  }
  if (bad) {
    snprintf(h->err, sizeof h->err, "error: too many neighbours %d %d", bad_ind, bad_n);
    return PFM_E_OVERFLOW;
  }
  return PFM_OK;
}

/* find_neighbour_distances (md_neighbours.f90:104-124) */
static void find_neighbour_distances(octx* h, onl* nl, const ogroup* g1, const ogroup* g2) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double dr2 = fdist(&h->pos[3 * g1->idx[i]], &h->pos[3 * g2->idx[nl->nlist[e]]], h->hbox,
                         &nl->dr[3 * e]);
      nl->moddr[e] = sqrt(dr2);
    }
  }
}

/* update_neighbour_list (md_neighbours.f90:32-52) */
static int update_neighbour_list(octx* h, int step, onl* nl, const ogroup* g1, const ogroup* g2) {
  if (g1->n > 0 && g2->n > 0 && nl->N > 0) {
    if (step % nl->period == 0) return find_neighbours(h, nl, g1, g2);
    find_neighbour_distances(h, nl, g1, g2);
  }
  return PFM_OK;
}

/* converce_neighbour_list (md_neighbours.f90:128-160), serial as the reference */
static int converse_list(octx* h, onl* cnl, const ogroup* g2, const onl* nl) {
  if (g2->n > 0 && nl->N > 0 && cnl->N > 0) {
    if (g2->n > cnl->N) return oerr(h, PFM_E_INVALID, "error: group2%N>cnl%N");
    for (int i = 0; i < g2->n; ++i) cnl->pidx[i] = g2->idx[i];
    for (int i = 0; i < cnl->N; ++i) cnl->nnum[i] = 0;
    for (int i = 0; i < nl->N; ++i) {
      for (int p = 0; p < nl->nnum[i]; ++p) {
        size_t e = (size_t)i * nl->nmax + p;
        int j = nl->nlist[e];
        cnl->nnum[j]++;
        if (cnl->nnum[j] > cnl->nmax) {
          snprintf(h->err, sizeof h->err, "error: too many neighbours %d %d", j + 1, cnl->nnum[j]);
          return PFM_E_OVERFLOW;
        }
        size_t c = (size_t)j * cnl->nmax + (cnl->nnum[j] - 1);
        cnl->nlist[c] = i;
        for (int k = 0; k < 3; ++k) cnl->dr[3 * c + k] = -nl->dr[3 * e + k];
        cnl->moddr[c] = nl->moddr[e];
      }
    }
  }
  return PFM_OK;
}

/* find_gr_nearest_neighbors (graphenenorm.f90:8-36) */
static int gr_nearest(octx* h, onl* nn, const onl* nl) {
  if (nn->nmax != 3) return oerr(h, PFM_E_GRAPHENE_NN, "error: nl_nn%neighb_num_max/=nnum_nn");
  for (int i = 0; i < nl->N; ++i) {
    int k = 0;
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      if (nl->moddr[e] < nn->rcut) {
        k++;
        if (k <= 3) {
          size_t c = (size_t)i * 3 + (k - 1);
          nn->nlist[c] = nl->nlist[e];
          nn->moddr[c] = nl->moddr[e];
          for (int d = 0; d < 3; ++d) nn->dr[3 * c + d] = nl->dr[3 * e + d];
        } else {
          snprintf(h->err, sizeof h->err, "error: too many gr nearest neibs %d %d", i + 1, p + 1);
          return PFM_E_GRAPHENE_NN;
        }
      }
    }
    if (k != 3) {
      snprintf(h->err, sizeof h->err, "error: not enough gr nearest neibs %d", i + 1);
      return PFM_E_GRAPHENE_NN;
    }
    nn->nnum[i] = 3;
  }
  return PFM_OK;
}

/* find_norm_in_graphene (graphenenorm.f90:38-56) */
static void gr_norm(double* gn, int n, const onl* nn) {
  for (int i = 0; i < n; ++i) {
    const double* d = &nn->dr[(size_t)i * 9];
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) { a[k] = d[3 + k] - d[k]; b[k] = d[k] - d[6 + k]; }
    double* g = &gn[3 * i];
    for (int k = 0; k < 3; ++k) {
      int k1 = (k + 1) % 3, k2 = (k + 2) % 3; /* mod(k,3)+1, mod(k+1,3)+1 in 1-based terms */
      g[k] = a[k1] * b[k2] - a[k2] * b[k1];
    }
    if (g[2] < 0.0) { g[0] = -g[0]; g[1] = -g[1]; g[2] = -g[2]; }
    double s = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    g[0] /= s; g[1] /= s; g[2] /= s;
  }
}

/* update_interactions_neighbour_lists (md_interactions.f90:138-178) */
static int update_lists(octx* h, int step) {
  int rc;
  for (int ii = 0; ii < h->ninter; ++ii) {
    ointer* it = &h->inter[ii];
    rc = update_neighbour_list(h, step, &it->nl[0], &h->groups[it->g[0]], &h->groups[it->g[1]]);
    if (rc) return rc;
    if (it->kind == PFM_LJ || it->kind == PFM_LJC || it->kind == PFM_MORSEC) {
      rc = converse_list(h, &it->nl[1], &h->groups[it->g[1]], &it->nl[0]);
      if (rc) return rc;
    }
    if (it->kind == PFM_LJC || it->kind == PFM_MORSEC) {
      int j;
      for (j = 0; j < h->ninter; ++j) if (h->inter[j].kind == PFM_TB || h->inter[j].kind == PFM_REBOSC) break;
      if (j < h->ninter) {
        /* update_nearest_neighbours_in_graphene (graphenenorm.f90:58-71) */
        onl* tbl = &h->inter[j].nl[0];
        const ogroup* gt = &h->groups[h->inter[j].g[0]];
        if (step % tbl->period == 0) {
          rc = gr_nearest(h, &it->nl[2], tbl);
          if (rc) return rc;
        } else {
          find_neighbour_distances(h, &it->nl[2], gt, gt);
        }
      } else {
        rc = update_neighbour_list(h, step, &it->nl[2], &h->groups[it->g[4]], &h->groups[it->g[5]]);
        if (rc) return rc;
      }
    }
  }
  /* update_norm_in_graphene (md_interactions.f90:195-208) */
  for (int ii = 0; ii < h->ninter; ++ii) {
    ointer* it = &h->inter[ii];
    if (it->kind == PFM_LJC || it->kind == PFM_MORSEC) gr_norm(it->gr_norm, it->nl[2].N, &it->nl[2]);
  }
  return PFM_OK;
}

/* -------------------------------------------------------------- potentials */

/* LJ_forces (LennardJones.f90:48-69) */
static void lj_forces(octx* h, const ointer* it, const onl* nl) {
  double eps = it->p[0], sig = it->p[1], R1 = it->p[2], R2 = it->p[3];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    double* f = &h->force[3 * nl->pidx[i]];
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double r = nl->moddr[e];
      if (r < R2) {
        double V = pow6(sig / r);
        double c = 4.0 * eps * (V * (12.0 * V - 6.0) / (r * r) * fcut(r, R1, R2) - V * (V - 1.0) * dfcut(r, R1, R2));
        for (int k = 0; k < 3; ++k) f[k] = f[k] - c * nl->dr[3 * e + k];
      }
    }
  }
}

/* LJ_energy (LennardJones.f90:23-46) */
static double lj_energy(const ointer* it, const onl* nl) {
  double eps = it->p[0], sig = it->p[1], R1 = it->p[2], R2 = it->p[3], E = 0.0;
  for (int i = 0; i < nl->N; ++i) {
    double ei = 0.0;
    for (int p = 0; p < nl->nnum[i]; ++p) {
      double r = nl->moddr[(size_t)i * nl->nmax + p];
      if (r < R2) { double V = pow6(sig / r); ei += 4 * eps * V * (V - 1.0) * fcut(r, R1, R2); }
    }
    E += ei;
  }
  return E;
}

/* scalar_lj_force (LennardJones_1g.f90:103-117) */
static inline double scalar_lj_force(double r, const ointer* it) {
  double invr2 = 1.0 / (r * r), U = invr2 * invr2 * invr2, f, dfr;
  fdfr_poly(r, it->p[2], it->p[3], &f, &dfr);
  return U * invr2 * ((it->c12t12 * U - it->c6t6) * f - (it->c12 * U - it->c6) * dfr);
}

/* LJ1g_forces (LennardJones_1g.f90:54-101): half list up to lessnnum,
 * Newton pairs; serial accumulation here. */
static void lj1g_forces(octx* h, const ointer* it, const onl* nl) {
  for (int i = 0; i < nl->N; ++i) {
    for (int p = 0; p < nl->lessnnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double F = scalar_lj_force(nl->moddr[e], it);
      int ind = nl->pidx[i], jnd = nl->pidx[nl->nlist[e]];
      for (int k = 0; k < 3; ++k) {
        double fp = F * nl->dr[3 * e + k];
        h->force[3 * ind + k] -= fp;
        h->force[3 * jnd + k] += fp;
      }
    }
  }
}

/* LJ1g_energy (LennardJones_1g.f90:28-52) */
static double lj1g_energy(const ointer* it, const onl* nl) {
  double E = 0.0;
  for (int i = 0; i < nl->N; ++i) {
    double ei = 0.0;
    for (int p = 0; p < nl->lessnnum[i]; ++p) {
      double r = nl->moddr[(size_t)i * nl->nmax + p];
      double U = 1.0 / (r * r * r * r * r * r);
      ei += U * (it->c12 * U - it->c6) * fcut_poly(r, it->p[2], it->p[3]);
    }
    E += ei;
  }
  return E;
}

/* Radial parts of ljc (LennardJonesCosine.f90) and morsec (MorseCosine.f90).
 * For the pair (r, dr, n):  E = pref*(V1 - m*V2*V3)*fc, with
 *   ljc:    V2 = (sig/r)^6, V1 = V2^2, pref = 4 eps, m = 1
 *   morsec: V2 = exp(-a(r-r0)), V1 = V2^2, pref = d, m = 2
 */
typedef struct { double pref, R1, R2, delt; int morse; double sig, a, r0; } cosprm;
static cosprm cos_params(const ointer* it) {
  cosprm c;
  memset(&c, 0, sizeof c);
  if (it->kind == PFM_LJC) {
    c.pref = 4.0 * it->p[0]; c.sig = it->p[1]; c.delt = it->p[2]; c.R1 = it->p[3]; c.R2 = it->p[4];
    c.morse = 0;
  } else {
    c.pref = it->p[0]; c.r0 = it->p[1]; c.a = it->p[2]; c.delt = it->p[3]; c.R1 = it->p[4]; c.R2 = it->p[5];
    c.morse = 1;
  }
  return c;
}
static inline double cos_V2(const cosprm* c, double r) {
  return c->morse ? exp(-c->a * (r - c->r0)) : pow6(c->sig / r);
}

/* MorseC_forces_for_graphene (MorseCosine.f90:53-110) and
 * LJC_forces_for_graphene (LennardJonesCosine.f90:53-110) */
static void cos_forces_graphene(octx* h, ointer* it) {
  cosprm c = cos_params(it);
  const onl* nl = &it->nl[0];
  const onl* nn = &it->nl[2];
  int nnum_nn = 3;
  double* gn = it->gr_norm;
  if ((int)it->p[it->kind == PFM_LJC ? 5 : 6]) { /* simplified */
    nnum_nn = 0;
    for (int i = 0; i < it->nl[2].N; ++i) { gn[3 * i] = 0; gn[3 * i + 1] = 0; gn[3 * i + 2] = 1.0; }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    double* f = &h->force[3 * nl->pidx[i]];
    const double* n = &gn[3 * i];
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double r = nl->moddr[e];
      if (r < c.R2) {
        const double* d = &nl->dr[3 * e];
        double V2 = cos_V2(&c, r), V1 = V2 * V2;
        double nd = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        double V3 = pow(fabs(nd) / r, c.delt);
        double fc = fcut(r, c.R1, c.R2), dfc = dfcut(r, c.R1, c.R2);
        double a, b;
        if (c.morse) {
          a = (2.0 * (c.a * V1 - (c.a + c.delt / r) * V2 * V3) / r * fc - (V1 - 2.0 * V2 * V3) * dfc);
          b = (2.0 * c.delt * V2 * V3 / nd * fc);
        } else {
          a = ((12.0 * V1 - (6.0 + c.delt) * V2 * V3) / (r * r) * fc - (V1 - V2 * V3) * dfc);
          b = (c.delt * V2 * V3 / nd * fc);
        }
        for (int k = 0; k < 3; ++k) f[k] = f[k] - c.pref * (a * d[k] + b * n[k]);
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    double* f = &h->force[3 * nl->pidx[i]];
    for (int q = 0; q < nnum_nn; ++q) {
      int j = nn->nlist[(size_t)i * 3 + q];
      int l1;
      for (l1 = 0; l1 < 3; ++l1) if (nn->nlist[(size_t)j * 3 + l1] == i) break;
      if (l1 > 2) {
#pragma omp critical
        snprintf(h->err, sizeof h->err, "l1>3");
        continue;
      }
      int l2 = (l1 + 1) % 3, l3 = (l1 + 2) % 3;
      const double* dj = &nn->dr[(size_t)j * 9];
      double d12[3], d31[3], d23[3];
      for (int k = 0; k < 3; ++k) {
        d12[k] = dj[3 * l2 + k] - dj[3 * l1 + k];
        d31[k] = dj[3 * l1 + k] - dj[3 * l3 + k];
        d23[k] = dj[3 * l3 + k] - dj[3 * l2 + k];
      }
      double s1212 = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
      double s3131 = d31[0] * d31[0] + d31[1] * d31[1] + d31[2] * d31[2];
      double s1231 = d12[0] * d31[0] + d12[1] * d31[1] + d12[2] * d31[2];
      double s2331 = d23[0] * d31[0] + d23[1] * d31[1] + d23[2] * d31[2];
      double s2312 = d23[0] * d12[0] + d23[1] * d12[1] + d23[2] * d12[2];
      double den = s1212 * s3131 - s1231 * s1231;
      const double* nj = &gn[3 * j];
      for (int p = 0; p < nl->nnum[j]; ++p) {
        size_t e = (size_t)j * nl->nmax + p;
        double r = nl->moddr[e];
        if (r < c.R2) {
          const double* d = &nl->dr[3 * e];
          double V2 = cos_V2(&c, r);
          double nd = nj[0] * d[0] + nj[1] * d[1] + nj[2] * d[2];
          double V3 = pow(fabs(nd) / r, c.delt);
          double fc = fcut(r, c.R1, c.R2);
          double w = 0.0;
          for (int k = 0; k < 3; ++k) w += d[k] * (d12[k] * s2331 - d31[k] * s2312);
          double coef = (c.morse ? 2.0 * c.pref : c.pref) * c.delt * V2 * V3 * fc / den / nd;
          for (int k = 0; k < 3; ++k) f[k] = f[k] - coef * nj[k] * w;
        }
      }
    }
  }
}

/* MorseC_forces_for_other_atoms (MorseCosine.f90:112-139),
 * LJC_forces_for_other_atoms (LennardJonesCosine.f90:112-139) */
static void cos_forces_other(octx* h, ointer* it) {
  cosprm c = cos_params(it);
  const onl* nl = &it->nl[1];
  const double* gn = it->gr_norm;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    double* f = &h->force[3 * nl->pidx[i]];
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double r = nl->moddr[e];
      if (r < c.R2) {
        const double* n = &gn[3 * nl->nlist[e]];
        const double* d = &nl->dr[3 * e];
        double V2 = cos_V2(&c, r), V1 = V2 * V2;
        double nd = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        double V3 = c.morse ? pow(fabs(nd) / (sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) * r), c.delt)
                            : pow(fabs(nd) / r, c.delt);
        double fc = fcut(r, c.R1, c.R2), dfc = dfcut(r, c.R1, c.R2);
        double a, b;
        if (c.morse) {
          a = (2.0 * (c.a * V1 - (c.a + c.delt / r) * V2 * V3) / r * fc - (V1 - 2.0 * V2 * V3) * dfc);
          b = (2.0 * c.delt * V2 * V3 / nd * fc);
        } else {
          a = ((12.0 * V1 - (6.0 + c.delt) * V2 * V3) / (r * r) * fc - (V1 - V2 * V3) * dfc);
          b = (c.delt * V2 * V3 / nd * fc);
        }
        for (int k = 0; k < 3; ++k) f[k] = f[k] - c.pref * (a * d[k] + b * n[k]);
      }
    }
  }
}

/* MorseC_energy (MorseCosine.f90:26-51), LJC_energy (LennardJonesCosine.f90:26-51) */
static double cos_energy(const ointer* it) {
  cosprm c = cos_params(it);
  const onl* nl = &it->nl[0];
  double E = 0.0;
  for (int i = 0; i < nl->N; ++i) {
    const double* n = &it->gr_norm[3 * i];
    double ei = 0.0;
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double r = nl->moddr[e];
      if (r < c.R2) {
        const double* d = &nl->dr[3 * e];
        double V2 = cos_V2(&c, r), V1 = V2 * V2;
        double V3 = pow(fabs(n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) / r, c.delt);
        ei += c.pref * (V1 - (c.morse ? 2.0 : 1.0) * V2 * V3) * fcut(r, c.R1, c.R2);
      }
    }
    E += ei;
  }
  return E;
}

/* RJL_forces (RosatoGuillopeLegrand.f90:51-94): exponent tables, Eb, gather */
static void rjl_forces(octx* h, const ointer* it, const onl* nl) {
  double A0 = it->p[0], xi = it->p[1], pp = it->p[2], q = it->p[3], r0 = it->p[4], R1 = it->p[5], R2 = it->p[6];
  size_t ne = (size_t)nl->N * nl->nmax;
  double* expp = (double*)malloc(sizeof(double) * (ne ? ne : 1));
  double* expq = (double*)malloc(sizeof(double) * (ne ? ne : 1));
  double* Eb = (double*)malloc(sizeof(double) * (nl->N ? nl->N : 1));
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (int i = 0; i < nl->N; ++i)
      for (int p = 0; p < nl->nnum[i]; ++p) {
        size_t e = (size_t)i * nl->nmax + p;
        double r = nl->moddr[e];
        if (r <= R2) {
          expq[e] = exp(-2.0 * q * (r / r0 - 1.0));
          expp[e] = exp(-pp * (r / r0 - 1.0));
        }
      }
#pragma omp for schedule(static)
    for (int i = 0; i < nl->N; ++i) {
      double s = 0.0;
      for (int p = 0; p < nl->nnum[i]; ++p) {
        size_t e = (size_t)i * nl->nmax + p;
        double r = nl->moddr[e];
        if (r <= R2) s = s + expq[e] * fcut(r, R1, R2);
      }
      Eb[i] = sqrt(s);
    }
#pragma omp for schedule(static)
    for (int i = 0; i < nl->N; ++i) {
      double* f = &h->force[3 * nl->pidx[i]];
      for (int p = 0; p < nl->nnum[i]; ++p) {
        size_t e = (size_t)i * nl->nmax + p;
        double r = nl->moddr[e];
        if (r <= R2) {
          double fc = fcut(r, R1, R2), dfc = dfcut(r, R1, R2);
          double c = (2.0 * A0 * (pp / r0 * fc - dfc * r) * expp[e] -
                      xi * (q / r0 * fc - dfc / 2.0 * r) * (1.0 / Eb[i] + 1.0 / Eb[nl->nlist[e]]) * expq[e]) / r;
          for (int k = 0; k < 3; ++k) f[k] = f[k] - c * nl->dr[3 * e + k];
        }
      }
    }
  }
  free(expp); free(expq); free(Eb);
}

/* RJL_energy (RosatoGuillopeLegrand.f90:23-49) */
static double rjl_energy(const ointer* it, const onl* nl) {
  double A0 = it->p[0], xi = it->p[1], pp = it->p[2], q = it->p[3], r0 = it->p[4], R1 = it->p[5], R2 = it->p[6];
  double E = 0.0;
  for (int i = 0; i < nl->N; ++i) {
    double Eb2 = 0.0, Er = 0.0;
    for (int p = 0; p < nl->nnum[i]; ++p) {
      double r = nl->moddr[(size_t)i * nl->nmax + p];
      if (r <= R2) {
        Eb2 = Eb2 + exp(-2.0 * q * (r / r0 - 1.0)) * fcut(r, R1, R2);
        Er = Er + exp(-pp * (r / r0 - 1.0)) * fcut(r, R1, R2);
      }
    }
    E += A0 * Er - xi * sqrt(Eb2);
  }
  return E;
}

/* bond order table of TB_forces / TB_energy (TersoffBrenner.f90:81-91,33-45) */
static void tb_bond_order(const ointer* it, const onl* nl, double* B) {
  double R1 = it->p[8], R2 = it->p[9], a0 = it->p[5], delt = it->p[4], c02 = it->c02, d02 = it->d02;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double b = 0.0;
      if (nl->moddr[e] < R2) {
        for (int q = 0; q < nl->nnum[i]; ++q) {
          size_t g = (size_t)i * nl->nmax + q;
          if (p != q && nl->moddr[g] < R2) {
            const double *dp = &nl->dr[3 * e], *dq = &nl->dr[3 * g];
            double cs = 1.0 + (dp[0] * dq[0] + dp[1] * dq[1] + dp[2] * dq[2]) / (nl->moddr[e] * nl->moddr[g]);
            b = b + fcut(nl->moddr[g], R1, R2) * (1.0 + c02 / d02 - c02 / (d02 + cs * cs));
          }
        }
        b = pow(1.0 + a0 * b, -delt);
      }
      B[e] = b;
    }
  }
}

static inline int find_in_row(const onl* nl, int j, int i) {
  int l;
  for (l = 0; l < nl->nnum[j]; ++l) if (nl->nlist[(size_t)j * nl->nmax + l] == i) break;
  return l;
}

/* TB_forces (TersoffBrenner.f90:74-150) */
static void tb_forces(octx* h, const ointer* it, const onl* nl) {
  double D = it->p[0], S = it->p[1], beta = it->p[2], r0 = it->p[3], delt = it->p[4], a0 = it->p[5];
  double R1 = it->p[8], R2 = it->p[9], c02 = it->c02, d02 = it->d02;
  size_t ne = (size_t)nl->N * nl->nmax;
  double* B = (double*)calloc(ne ? ne : 1, sizeof(double));
  tb_bond_order(it, nl, B);
  double ex = 1.0 / delt + 1.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nl->N; ++i) {
    double* f = &h->force[3 * nl->pidx[i]];
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t ep = (size_t)i * nl->nmax + p;
      const double* drp = &nl->dr[3 * ep];
      double rp = nl->moddr[ep];
      double dB[3] = {0.0, 0.0, 0.0};
      for (int q = 0; q < nl->nnum[i]; ++q) {
        if (p == q) continue;
        size_t eq = (size_t)i * nl->nmax + q;
        const double* drq = &nl->dr[3 * eq];
        double rq = nl->moddr[eq];
        double rr = 1.0 / rp / rq;
        double cosi = (drp[0] * drq[0] + drp[1] * drq[1] + drp[2] * drq[2]) * rr;
        double t = d02 + (1.0 + cosi) * (1.0 + cosi);
        double c1 = fcut(rq, R1, R2) * 2.0 * a0 * c02 * (1.0 + cosi) / (t * t);
        double c2 = dfcut(rq, R1, R2) * a0 * (1.0 + c02 / d02 - c02 / t);
        for (int k = 0; k < 3; ++k)
          dB[k] = dB[k] + c1 * ((drp[k] + drq[k]) * rr - cosi * (drp[k] / (rp * rp) + drq[k] / (rq * rq))) + c2 * drq[k];
      }
      double bp = pow(B[ep], ex);
      for (int k = 0; k < 3; ++k) dB[k] = dB[k] * bp;
      int j = nl->nlist[ep];
      int l = find_in_row(nl, j, i);
      size_t el = (size_t)j * nl->nmax + l;
      const double* drl = &nl->dr[3 * el];
      double rl = nl->moddr[el];
      double bl = pow(B[el], ex);
      for (int q = 0; q < nl->nnum[j]; ++q) {
        if (q == l) continue;
        size_t eq = (size_t)j * nl->nmax + q;
        const double* drq = &nl->dr[3 * eq];
        double rq = nl->moddr[eq];
        double rr = 1.0 / rl / rq;
        double cosi = (drl[0] * drq[0] + drl[1] * drq[1] + drl[2] * drq[2]) * rr;
        double t = d02 + (1.0 + cosi) * (1.0 + cosi);
        double c1 = bl * fcut(rq, R1, R2) * 2.0 * a0 * c02 * (1.0 + cosi) / (t * t);
        for (int k = 0; k < 3; ++k) dB[k] = dB[k] + c1 * (-drq[k] * rr + cosi * drl[k] / (rl * rl));
      }
      for (int k = 0; k < 3; ++k) dB[k] = dB[k] * (-delt) / 2.0;
      if (rp < R2) {
        double fc = fcut(rp, R1, R2);
        double dff = dfcut(rp, R1, R2) / fc;
        double a = -sqrt(2.0 * S) * beta * (rp - r0);
        double ea = exp(a), eas = exp(a / S);
        double pre = fc * D / (S - 1.0);
        for (int k = 0; k < 3; ++k)
          f[k] = f[k] + pre * ((drp[k] * dff - sqrt(2.0 * S) * beta / rp * drp[k]) * ea -
                               (dB[k] + (B[ep] + B[el]) / 2 * (drp[k] * dff - sqrt(2.0 / S) * beta / rp * drp[k])) * S * eas);
      }
      for (int q = 0; q < nl->nnum[j]; ++q) {
        if (q == l) continue;
        size_t eq = (size_t)j * nl->nmax + q;
        const double* drq = &nl->dr[3 * eq];
        double rq = nl->moddr[eq];
        double fc = fcut(rl, R1, R2), dfc = dfcut(rl, R1, R2);
        double rr = 1.0 / rl / rq;
        double cosi = (drl[0] * drq[0] + drl[1] * drq[1] + drl[2] * drq[2]) * rr;
        double t = d02 + (1.0 + cosi) * (1.0 + cosi);
        double g1 = fc * 2.0 * a0 * c02 * (1.0 + cosi) / (t * t);
        double g2 = dfc * a0 * (1.0 + c02 / d02 - c02 / t);
        double tail = fcut(rq, R1, R2) * D / (S - 1.0) * S * exp(-sqrt(2.0 * S) * beta * (rq - r0) / S);
        double bq = delt / 2 * pow(B[eq], ex);
        for (int k = 0; k < 3; ++k)
          f[k] = f[k] + bq * (g1 * (-drq[k] * rr + cosi * drl[k] / (rl * rl)) + drp[k] * g2) * tail;
      }
    }
  }
  free(B);
}

/* TB_energy (TersoffBrenner.f90:26-72) */
static double tb_energy(const ointer* it, const onl* nl) {
  double D = it->p[0], S = it->p[1], beta = it->p[2], r0 = it->p[3], R1 = it->p[8], R2 = it->p[9];
  size_t ne = (size_t)nl->N * nl->nmax;
  double* B = (double*)calloc(ne ? ne : 1, sizeof(double));
  tb_bond_order(it, nl, B);
  double E = 0.0;
  for (int i = 0; i < nl->N; ++i) {
    double ei = 0.0;
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t ep = (size_t)i * nl->nmax + p;
      if (nl->moddr[ep] < R2) {
        int j = nl->nlist[ep];
        if (j > i) {
          int q = find_in_row(nl, j, i);
          double a = -sqrt(2.0 * S) * beta * (nl->moddr[ep] - r0);
          ei += fcut(nl->moddr[ep], R1, R2) * D / (S - 1.0) *
                (exp(a) - (B[ep] + B[(size_t)j * nl->nmax + q]) / 2 * S * exp(a / S));
        }
      }
    }
    E += ei;
  }
  free(B);
  return E;
}


/* ------------------------------------------------------------- rebosc */

/* REBOsc_energy (REBOsolidcarbon.f90:27-92).  Parameters in read order
 * (REBOsolidcarbon.f90:17-22): A Q alpha B(3) beta(3) T g(6) R1 R2. */
static double rebo_energy(const ointer* it, const onl* nl) {
  const double *P = it->p, A = P[0], Q = P[1], al = P[2], *B = P + 3, *be = P + 6, T = P[9], *g = P + 10;
  const double R1 = P[16], R2 = P[17];
  size_t ne = (size_t)nl->N * nl->nmax;
  double* bsp = (double*)calloc(ne ? ne : 1, sizeof(double));
  double* bdh = (double*)calloc(ne ? ne : 1, sizeof(double));
  for (int i = 0; i < nl->N; ++i) {
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      if (nl->moddr[e] < R2) {
        int j = nl->nlist[e];
        const double* dp = &nl->dr[3 * e];
        for (int q = 0; q < nl->nnum[i]; ++q) {
          size_t eq = (size_t)i * nl->nmax + q;
          if (p != q && nl->moddr[eq] < R2) {
            const double* dq = &nl->dr[3 * eq];
            double c = (dp[0] * dq[0] + dp[1] * dq[1] + dp[2] * dq[2]) / (nl->moddr[e] * nl->moddr[eq]);
            bsp[e] = bsp[e] + fcut(nl->moddr[eq], R1, R2) *
                                  (g[0] + g[1] * c + g[2] * c * c + g[3] * c * c * c + g[4] * c * c * c * c +
                                   g[5] * c * c * c * c * c);
            for (int l = 0; l < nl->nnum[j]; ++l) {
              size_t el = (size_t)j * nl->nmax + l;
              if (nl->nlist[el] != i && nl->moddr[el] < R2) {
                const double* dl = &nl->dr[3 * el];
                double aa = nl->moddr[e] * nl->moddr[e];
                double ab = dp[0] * dq[0] + dp[1] * dq[1] + dp[2] * dq[2];
                double ac = dp[0] * dl[0] + dp[1] * dl[1] + dp[2] * dl[2];
                double bc = dq[0] * dl[0] + dq[1] * dl[1] + dq[2] * dl[2];
                double x = aa * bc - ab * ac;
                bdh[e] = bdh[e] + fcut(nl->moddr[eq], R1, R2) * fcut(nl->moddr[el], R1, R2) *
                                      (1. - x * x / (aa * nl->moddr[eq] * nl->moddr[eq] - ab * ab) /
                                                (aa * nl->moddr[el] * nl->moddr[el] - ac * ac));
              }
            }
          }
        }
        bsp[e] = pow(1. + bsp[e], -0.5);
        bdh[e] = T * bdh[e];
      }
    }
  }
  double E = 0.0;
  for (int i = 0; i < nl->N; ++i) {
    for (int p = 0; p < nl->nnum[i]; ++p) {
      size_t e = (size_t)i * nl->nmax + p;
      double r = nl->moddr[e];
      if (r < R2) {
        int j = nl->nlist[e];
        if (j > i) {
          int q = find_in_row(nl, j, i);
          double bq = q < nl->nmax ? bsp[(size_t)j * nl->nmax + q] : 0.0;
          E = E + fcut(r, R1, R2) * ((1 + Q / r) * A * exp(-al * r) -
                                     ((bsp[e] + bq) / 2 + bdh[e]) *
                                         (B[0] * exp(-be[0] * r) + B[1] * exp(-be[1] * r) + B[2] * exp(-be[2] * r)));
        }
      }
    }
  }
  free(bsp); free(bdh);
  return E;
}

/* calculate_truncated_nl (md_interactions.f90:337-384) for row i with
 * neighbour order n: the rows of the atoms within n-1 hops of i copied
 * whole, the atoms at n hops given their first back entry. */
static void truncated_nl(onl* t, const onl* nl, int i, int n) {
  for (int j = 0; j < t->N; ++j) { t->pidx[j] = -1; t->nnum[j] = 0; }
  size_t m = nl->nmax;
  memcpy(&t->nlist[i * m], &nl->nlist[i * m], sizeof(int) * nl->nnum[i]);
  memcpy(&t->dr[3 * i * m], &nl->dr[3 * i * m], sizeof(double) * 3 * nl->nnum[i]);
  memcpy(&t->moddr[i * m], &nl->moddr[i * m], sizeof(double) * nl->nnum[i]);
  t->nnum[i] = nl->nnum[i];
  t->pidx[i] = nl->pidx[i];
  for (int ni = 1; ni <= n; ++ni) {
    if (ni < n) {
      for (int j = 0; j < t->N; ++j)
        for (int p = 0; p < t->nnum[j]; ++p) {
          int q = t->nlist[j * m + p];
          if (t->nnum[q] == 0) {
            memcpy(&t->nlist[q * m], &nl->nlist[q * m], sizeof(int) * nl->nnum[q]);
            memcpy(&t->dr[3 * q * m], &nl->dr[3 * q * m], sizeof(double) * 3 * nl->nnum[q]);
            memcpy(&t->moddr[q * m], &nl->moddr[q * m], sizeof(double) * nl->nnum[q]);
            t->pidx[q] = nl->pidx[q];
          }
        }
      for (int j = 0; j < t->N; ++j)
        if (t->pidx[j] >= 0) t->nnum[j] = nl->nnum[j];
    } else {
      for (int j = 0; j < t->N; ++j)
        for (int p = 0; p < t->nnum[j]; ++p) {
          int q = t->nlist[j * m + p];
          if (t->nnum[q] == 0) {
            t->nnum[q] = t->nnum[q] + 1;
            size_t e = q * m + (t->nnum[q] - 1);
            t->nlist[e] = j;
            for (int k = 0; k < 3; ++k) t->dr[3 * e + k] = -nl->dr[3 * (j * m + p) + k];
            t->moddr[e] = nl->moddr[j * m + p];
            t->pidx[q] = nl->pidx[q];
          }
        }
    }
  }
}

/* shift_drs (md_interactions.f90:386-405) for a list of one row set (nl_n = 1) */
static void shift_drs(onl* t, int inl, int k, double dx) {
  size_t m = t->nmax;
  for (int p = 0; p < t->nnum[inl]; ++p) {
    double* d = &t->dr[3 * (inl * m + p)];
    d[k] = d[k] - dx;
    t->moddr[inl * m + p] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  }
  for (int j = 0; j < t->N; ++j)
    for (int p = 0; p < t->nnum[j]; ++p)
      if (t->nlist[j * m + p] == inl) {
        double* d = &t->dr[3 * (j * m + p)];
        d[k] = d[k] + dx;
        t->moddr[j * m + p] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      }
}

/* calculate_forces_numerically (md_interactions.f90:273-310) for rebosc
 * (neighbour order 3, md_interactions.f90:111): central differences of the
 * energy of the truncated list, dx = 1e-6 A. */
static void rebo_forces(octx* h, const ointer* it, const onl* nl) {
  const double dx = 1e-6;
  onl t;
  memset(&t, 0, sizeof t);
  t.N = nl->N; t.nmax = nl->nmax;
  nl_alloc(&t);
  for (int inl = 0; inl < nl->N; ++inl) {
    truncated_nl(&t, nl, inl, 3);
    for (int k = 0; k < 3; ++k) {
      shift_drs(&t, inl, k, -dx);
      double e1 = rebo_energy(it, &t);
      shift_drs(&t, inl, k, 2 * dx);
      double e2 = rebo_energy(it, &t);
      if (k != 2) shift_drs(&t, inl, k, -dx);
      h->force[3 * nl->pidx[inl] + k] += (e1 - e2) / 2 / dx;
    }
  }
  free(t.nlist); free(t.nnum); free(t.lessnnum); free(t.pidx); free(t.dr); free(t.moddr);
}

/* calculate_forces (md_interactions.f90:210-242) */
static int calc_forces(octx* h) {
  for (int ii = 0; ii < h->ninter; ++ii) {
    ointer* it = &h->inter[ii];
    switch (it->kind) {
      case PFM_LJ: lj_forces(h, it, &it->nl[0]); lj_forces(h, it, &it->nl[1]); break;
      case PFM_LJ1G: lj1g_forces(h, it, &it->nl[0]); break;
      case PFM_LJC:
      case PFM_MORSEC:
        h->err[0] = 0;
        cos_forces_graphene(h, it);
        if (h->err[0]) return PFM_E_GRAPHENE_NN;
        cos_forces_other(h, it);
        break;
      case PFM_TB: tb_forces(h, it, &it->nl[0]); break;
      case PFM_RJL: rjl_forces(h, it, &it->nl[0]); break;
    }
  }
  /* calculate_forces_numerically (md_simulation.f90:168) */
  for (int ii = 0; ii < h->ninter; ++ii)
    if (h->inter[ii].kind == PFM_REBOSC) rebo_forces(h, &h->inter[ii], &h->inter[ii].nl[0]);
  return PFM_OK;
}

/* calculate_potential_energies (md_interactions.f90:244-271) */
static void calc_energies(octx* h) {
  double pe = 0.0;
  for (int ii = 0; ii < h->ninter; ++ii) {
    ointer* it = &h->inter[ii];
    switch (it->kind) {
      case PFM_LJ: it->energy = lj_energy(it, &it->nl[0]); break;
      case PFM_LJ1G: it->energy = lj1g_energy(it, &it->nl[0]); break;
      case PFM_LJC: case PFM_MORSEC: it->energy = cos_energy(it); break;
      case PFM_TB: it->energy = tb_energy(it, &it->nl[0]); break;
      case PFM_RJL: it->energy = rjl_energy(it, &it->nl[0]); break;
      case PFM_REBOSC: it->energy = rebo_energy(it, &it->nl[0]); break;
    }
    pe += it->energy;
  }
  h->pe = pe;
}

/* ------------------------------------------------ integrators and reductions */

static const ogroup* grp(octx* h, int g) { return &h->groups[g]; }

/* calculate_kinetic_energy (md_general.f90:161-182) */
static double kinetic_energy(octx* h, const ogroup* g) {
  double ke = 0.0;
  for (int ind = 0; ind < g->n; ++ind) {
    int i = g->idx[ind];
    const double* v = &h->vel[3 * i];
    ke = ke + h->mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / 2 * MASS_COEF;
  }
  return ke;
}

/* zero_momentum (md_general.f90:236-253) via calculate_mass_center_velocity (:210-234) */
static void zero_momentum(octx* h, const ogroup* g) {
  double mcv[3] = {0, 0, 0}, totm = 0.0;
  for (int ind = 0; ind < g->n; ++ind) {
    int i = g->idx[ind];
    for (int k = 0; k < 3; ++k) mcv[k] += h->mass[i] * h->vel[3 * i + k];
  }
  for (int ind = 0; ind < g->n; ++ind) totm += h->mass[g->idx[ind]];
  for (int k = 0; k < 3; ++k) mcv[k] = mcv[k] / totm;
  for (int ind = 0; ind < g->n; ++ind)
    for (int k = 0; k < 3; ++k) h->vel[3 * g->idx[ind] + k] -= mcv[k];
}

/* integrate_verlet_xyz/z_velocities (md_integrators.f90:58-97) */
static void kick(octx* h, const ogroup* g, int only_z) {
  for (int ind = 0; ind < g->n; ++ind) {
    int i = g->idx[ind];
    for (int k = only_z ? 2 : 0; k < 3; ++k)
      h->vel[3 * i + k] = h->vel[3 * i + k] + h->force[3 * i + k] / h->mass[i] / MASS_COEF * h->ts[1];
  }
}
/* integrate_verlet_xyz/z_positions (md_integrators.f90:7-56) */
static void drift(octx* h, const ogroup* g, int only_z) {
  for (int ind = 0; ind < g->n; ++ind) {
    int i = g->idx[ind];
    for (int k = only_z ? 2 : 0; k < 3; ++k) {
      double x = h->pos[3 * i + k] + h->vel[3 * i + k] * h->ts[0];
      if (x > h->box[k]) x = x - h->box[k];
      else if (x < 0.0) x = x + h->box[k];
      h->pos[3 * i + k] = x;
    }
  }
}
/* molecular_static_xyz_velocities / _1D_velocities (md_integrators.f90:99-145) */
static void ms_xyz(octx* h, const ogroup* g) {
  for (int ind = 0; ind < g->n; ++ind) {
    int i = g->idx[ind];
    double *v = &h->vel[3 * i], *f = &h->force[3 * i];
    double fv = f[0] * v[0] + f[1] * v[1] + f[2] * v[2];
    double ff = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    if (fv > 0.0 && ff > 1e-12) {
      for (int k = 0; k < 3; ++k) v[k] = fv / ff * f[k];
    } else {
      v[0] = v[1] = v[2] = 0.0;
    }
  }
}
static void ms_1d(octx* h, const ogroup* g) {
  for (int ind = 0; ind < g->n; ++ind) {
    int i = g->idx[ind];
    double *v = &h->vel[3 * i], *f = &h->force[3 * i];
    double fv = f[0] * v[0] + f[1] * v[1] + f[2] * v[2];
    if (!(fv > 0.0)) v[0] = v[1] = v[2] = 0.0;
  }
}

/* integrate_nose_hoover_chain (md_integrators.f90:200-245) */
static void nhc_half(octx* h, onhc* c) {
  const ogroup* g = grp(h, c->group);
  double ke = kinetic_energy(h, g);
  double kt = KT_A_DEGREE * c->temperature;
  double kedif = 2.0 * ke - 3.0 * c->L * kt;
  int M = c->M;
  double b = 1.0;
  double *v = c->v, *q = c->q;
  if (M == 1) {
    v[0] = v[0] + kedif / q[0] * h->ts[2];
  } else {
    v[M - 1] = v[M - 1] + (q[M - 2] * v[M - 2] * v[M - 2] - kt) / q[M - 1] * h->ts[2];
    for (int i = M - 2; i >= 1; --i) {
      b = exp(-v[i + 1] * h->ts[3]);
      v[i] = v[i] * (b * b) + (q[i - 1] * v[i - 1] * v[i - 1] - kt) / q[i] * h->ts[2] * b;
    }
    b = exp(-v[1] * h->ts[3]);
    v[0] = v[0] * (b * b) + kedif / q[0] * h->ts[2] * b;
  }
  c->s = exp(-v[0] * h->ts[1]);
  for (int ind = 0; ind < g->n; ++ind)
    for (int k = 0; k < 3; ++k) h->vel[3 * g->idx[ind] + k] *= c->s;
  kedif = 2.0 * ke * (c->s * c->s) - 3.0 * c->L * kt;
  for (int i = 0; i < M; ++i) c->x[i] = c->x[i] + v[i] * h->ts[1];
  if (M == 1) {
    v[0] = v[0] + kedif / q[0] * h->ts[2];
  } else {
    v[0] = v[0] * (b * b) + kedif / q[0] * h->ts[2] * b;
    for (int i = 1; i <= M - 2; ++i) {
      b = exp(-v[i + 1] * h->ts[3]);
      v[i] = v[i] * (b * b) + (q[i - 1] * v[i - 1] * v[i - 1] - kt) / q[i] * h->ts[2] * b;
    }
    v[M - 1] = v[M - 1] + (q[M - 2] * v[M - 2] * v[M - 2] - kt) / q[M - 1] * h->ts[2];
  }
}

/* calculate_nose_hoover_chain_energy (md_integrators.f90:247-260) */
static void nhc_energy(onhc* c) {
  double kt = KT_A_DEGREE * c->temperature;
  c->e = c->q[0] / 2 * c->v[0] * c->v[0] + 3.0 * c->L * kt * c->x[0];
  for (int i = 1; i < c->M; ++i) c->e = c->e + c->q[i] / 2 * c->v[i] * c->v[i] + kt * c->x[i];
}

/* check_positions (md_general.f90:342-364) */
static int check_positions(octx* h, int step) {
  const double tol = 0.0000001;
  for (int i = 0; i < h->N; ++i)
    for (int k = 0; k < 3; ++k) {
      double x = h->pos[3 * i + k];
      if (!(x > (0.0 - tol) && x < (h->box[k] + tol))) {
        snprintf(h->err, sizeof h->err, "%d particle out of cell %.10g %.10g %.10g (step %d)", i + 1,
                 h->pos[3 * i], h->pos[3 * i + 1], h->pos[3 * i + 2], step);
        return PFM_E_OUT_OF_CELL;
      }
    }
  return PFM_OK;
}

/* invert_z_velocities (md_general.f90:382-398) with md()'s borders (md_simulation.f90:140) */
static void invert_z(octx* h) {
  double lo = 0.8 * h->box[2], hi = 0.9 * h->box[2], mid = (lo + hi) / 2;
  for (int i = 0; i < h->N; ++i) {
    double z = h->pos[3 * i + 2], vz = h->vel[3 * i + 2];
    if ((z > lo && z < mid && vz > 0.0) || (z < hi && z > mid && vz < 0.0)) h->vel[3 * i + 2] = -vz;
  }
}

/* --------------------------------------------------------------- the step */

/* The body of md()'s step loop (md_simulation.f90:137-231), for steps
 * step_begin..step_end of one integrator.  init_time_steps
 * (md_general.f90:45-55) sets ts(i) = dt/2**(i-1). */
int orc_run(octx* h, int integ, double dt, int s0, int s1, int energy_at_end, double ms_de,
            int* last_step, int* event) {
  if (integ < PFM_NVT || integ > PFM_NVMS) return oerr(h, PFM_E_INVALID, "error: unknown integrator");
  for (int i = 0; i < 4; ++i) h->ts[i] = dt / (double)(1 << i);
  *event = PFM_EV_DONE;
  int rc;
  for (int s = s0; s <= s1; ++s) {
    *last_step = s;
    for (int c = 0; c < h->nchange; ++c)
      change_group_n(&h->groups[h->change[c].to], s, &h->change[c], &h->groups[h->change[c].from]);
    rc = check_positions(h, s);
    if (rc) return rc;
    if (h->invert_z) invert_z(h);
    if (s != 0) {
      if (integ == PFM_NVMS && fabs(h->pe - h->pe_prev) <= ms_de) { *event = PFM_EV_CONVERGED; return PFM_OK; }
      if (integ == PFM_NVT) for (int c = 0; c < h->nnhc; ++c) nhc_half(h, &h->nhc[c]);
      kick(h, grp(h, h->g_xyz), 0);
      kick(h, grp(h, h->g_z), 1);
      drift(h, grp(h, h->g_xyz), 0);
      drift(h, grp(h, h->g_z), 1);
    }
    h->pe_prev = h->pe;
    rc = update_lists(h, s);
    if (rc) return rc;
    if (s % h->zm_period == 0) zero_momentum(h, grp(h, h->g_all));
    { /* zero_forces (md_integrators.f90:147-163) */
      const ogroup* g = grp(h, h->g_all);
      for (int ind = 0; ind < g->n; ++ind)
        for (int k = 0; k < 3; ++k) h->force[3 * g->idx[ind] + k] = 0.0;
    }
    rc = calc_forces(h);
    if (rc) return rc;
    if (s != 0) {
      kick(h, grp(h, h->g_xyz), 0);
      kick(h, grp(h, h->g_z), 1);
      if (integ == PFM_NVT) for (int c = 0; c < h->nnhc; ++c) nhc_half(h, &h->nhc[c]);
      if (integ == PFM_NVMS) { ms_xyz(h, grp(h, h->g_xyz)); ms_1d(h, grp(h, h->g_z)); }
      h->sim_time = h->sim_time + h->ts[0];
    }
    if ((s == s1 && energy_at_end) || integ == PFM_NVMS) {
      calc_energies(h);
      /* calculate_temperature (md_general.f90:301-311) */
      const ogroup* g = grp(h, h->g_all_moving);
      h->ke = kinetic_energy(h, g);
      h->temp = 2 * h->ke / KT_A_DEGREE / (3 * g->n);
      if (integ == PFM_NVT) for (int c = 0; c < h->nnhc; ++c) nhc_energy(&h->nhc[c]);
    }
  }
  return PFM_OK;
}

int orc_energies(octx* h, double* e_inter, double* pe, double* ke, double* temp, double* e_nhc,
                 double* sim_time) {
  if (e_inter) for (int i = 0; i < h->ninter; ++i) e_inter[i] = h->inter[i].energy;
  if (pe) *pe = h->pe;
  if (ke) *ke = h->ke;
  if (temp) *temp = h->temp;
  if (e_nhc) for (int i = 0; i < h->nnhc; ++i) e_nhc[i] = h->nhc[i].e;
  if (sim_time) *sim_time = h->sim_time;
  return PFM_OK;
}

int orc_get_state(octx* h, double* pos, double* vel, double* force) {
  if (pos) memcpy(pos, h->pos, sizeof(double) * 3 * h->N);
  if (vel) memcpy(vel, h->vel, sizeof(double) * 3 * h->N);
  if (force) memcpy(force, h->force, sizeof(double) * 3 * h->N);
  return PFM_OK;
}

int orc_set_state(octx* h, const double* pos, const double* vel) {
  if (pos) memcpy(h->pos, pos, sizeof(double) * 3 * h->N);
  if (vel) memcpy(h->vel, vel, sizeof(double) * 3 * h->N);
  return PFM_OK;
}

int orc_get_list(octx* h, int inter, int k, int* n_rows, int* nnum, int* nlist, int* lessnnum,
                 int stride) {
  if (inter < 1 || inter > h->ninter) return oerr(h, PFM_E_INVALID, "error: bad interaction index");
  ointer* it = &h->inter[inter - 1];
  if (k < 1 || k > it->nl_n) return oerr(h, PFM_E_INVALID, "error: bad list index");
  onl* nl = &it->nl[k - 1];
  *n_rows = nl->N;
  if (nnum == NULL) return PFM_OK;
  for (int i = 0; i < nl->N; ++i) {
    nnum[i] = nl->nnum[i];
    if (lessnnum) lessnnum[i] = nl->lessnnum[i];
    if (nl->nnum[i] > stride) return oerr(h, PFM_E_INVALID, "error: stride smaller than a row");
    for (int p = 0; p < nl->nnum[i]; ++p) nlist[(size_t)i * stride + p] = nl->nlist[(size_t)i * nl->nmax + p] + 1;
  }
  return PFM_OK;
}

int orc_get_normals(octx* h, int inter, double* out) {
  if (inter < 1 || inter > h->ninter) return oerr(h, PFM_E_INVALID, "error: bad interaction index");
  ointer* it = &h->inter[inter - 1];
  if (!it->gr_norm) return oerr(h, PFM_E_INVALID, "error: interaction has no graphene normals");
  memcpy(out, it->gr_norm, sizeof(double) * 3 * it->nl[2].N);
  return PFM_OK;
}

int orc_last_error(octx* h, char* buf, int len) {
  if (len > 0) { strncpy(buf, h->err, len - 1); buf[len - 1] = 0; }
  return PFM_OK;
}

void orc_destroy(octx* h) {
  if (!h) return;
  for (int i = 0; i < h->ninter; ++i) {
    ointer* it = &h->inter[i];
    for (int j = 0; j < it->nl_n; ++j) {
      onl* nl = &it->nl[j];
      free(nl->nlist); free(nl->nnum); free(nl->lessnnum); free(nl->pidx); free(nl->dr); free(nl->moddr);
    }
    free(it->gr_norm);
  }
  for (int g = 0; g <= MAXG; ++g) free(h->groups[g].idx);
  free(h->pos); free(h->vel); free(h->mass); free(h->force);
  free(h);
}
