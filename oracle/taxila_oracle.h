/*
 * taxila_oracle.h -- CPU restatement of the Taxila LBM flow step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP product
 * path and the CPU baseline of the benchmark; nothing in the product links
 * it, and only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * It keeps the reference's data layout (ghosted AoS arrays, 1-based owned
 * index space with PETSc-style ghost layers) and evaluation order so that it
 * reproduces the reference's FP64 arithmetic operation by operation; each
 * routine in taxila_oracle.c cites the reference lines it restates.
 */
#ifndef TAXILA_ORACLE_H
#define TAXILA_ORACLE_H

#include "taxila_lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

orc_t* orc_create(const txg_desc* d, char* err, int errlen);
void orc_destroy(orc_t* o);
int orc_set_walls(orc_t* o, const double* walls);          /* owned, natural */
int orc_set_state(orc_t* o, const double* rho, const double* u);
int orc_set_fi(orc_t* o, const double* fi);
int orc_set_bc_values(orc_t* o, int face, const double* vals); /* [(i*D+d)*C+m] */
int orc_init(orc_t* o);                                      /* LBMInit2 flow part */
int orc_step(orc_t* o, int nsteps);                          /* LBMRun2 flow body */
int orc_get_fi(const orc_t* o, double* fi);
int orc_get_moments(const orc_t* o, double* rho, double* flux, double* forces);
int orc_diagnostics(orc_t* o, double* rhot, double* prs, double* velt);
int orc_delta_norm(orc_t* o, double* norm);
int orc_set_threads(int n);
int orc_phase(orc_t* o, int phase);
/* Fluid-fluid stencil table in accumulation order: for entry e, off[3*e..],
 * ffw index k, and the wall-test alternatives (see taxila_oracle.c). */
int orc_ff_table(int dims, int isotropy, int* off, int* kidx, int* nalt,
                 int* altlen, int* altcells, int maxent);

#ifdef __cplusplus
}
#endif
#endif
