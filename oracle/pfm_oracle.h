/*
 * pfm_oracle.h -- CPU FP64 restatement of the PFMDS molecular-dynamics step
 * (reference: code_source/MOLECULAR_DYNAMICS and code_source/INTERACTION_POTENTIALS).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libpfmds.so and the Fortran front end) never links
 * or calls it.
 *
 * The entry points mirror the product C-ABI in include/pfm.h one for one
 * (orc_* <-> pfm_*), so a test drives both with identical calls.  Indices
 * crossing this boundary are 1-based, as in the reference.
 */
#ifndef PFM_ORACLE_H
#define PFM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct octx octx;

int  orc_create(octx** h, int n_atoms, const double box[3], const double* pos,
                const double* vel, const double* mass);
int  orc_set_group(octx* h, int gid, int n, const int* idx);
int  orc_set_roles(octx* h, int all_moving, int xyz_moving, int z_moving,
                   int all_atoms, int zero_momentum_period, int invert_z);
int  orc_add_interaction(octx* h, int kind, const double* params, int n_params,
                         int nl_n, const int* group_nums, const int* nmax,
                         const double* r_cut, const int* period);
int  orc_add_nhc(octx* h, int group, double temperature, int M, double q1);
int  orc_run(octx* h, int integrator, double dt, int step_begin, int step_end,
             int energy_at_end, double ms_de, int* last_step, int* event);
int  orc_energies(octx* h, double* e_inter, double* pe, double* ke, double* temp,
                  double* e_nhc, double* sim_time);
int  orc_get_state(octx* h, double* pos, double* vel, double* force);
int  orc_set_state(octx* h, const double* pos, const double* vel);
int  orc_get_list(octx* h, int inter, int k, int* n_rows, int* nnum, int* nlist,
                  int* lessnnum, int stride);
int  orc_get_normals(octx* h, int inter, double* gr_norm);
int  orc_get_chain(octx* h, int c, double* x, double* v);
int  orc_add_group_change(octx* h, int from, int to, int ts1, int ts2, int frec);
int  orc_group_count(octx* h, int gid, int* n);
int  orc_pe_change(octx* h, double* de);
int  orc_last_error(octx* h, char* buf, int len);
void orc_destroy(octx* h);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
