/* smaug_oracle.h -- CPU restatement of the SMAUG step (test infrastructure,
 * see smaug_oracle.c).  Shares the parameter struct of the product C-ABI. */
#ifndef SMAUG_ORACLE_H_
#define SMAUG_ORACLE_H_
#include "../include/smaug_hip.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef struct orc orc_t;
int orc_create(const smaug_params* p, const double* w0, const double* pos, orc_t** out);
int orc_step(orc_t* o, int nsteps);
int orc_get_state(orc_t* o, double* w);
int orc_set_state(orc_t* o, const double* w);
int orc_get_wd(orc_t* o, double* wd);
int orc_cfl(orc_t* o, double* cmax_per_dim, double* dt);
int orc_get_clock(orc_t* o, int* it, double* t, double* qt, double* dt);
int orc_probe_grad(orc_t* o, int kind, const double* a, int bcidx, int dir, double* out);
int orc_probe_visc(orc_t* o, int field, int dim, double* nur, double* nul, double* cmax);
int orc_restart(orc_t* o, const double* w, int it, double qt);
int orc_set_driver_map(orc_t* o, const double* amap);
int orc_set_driver_param(orc_t* o, int q, double v);
long orc_nu_mode(orc_t* o, int mode);
int orc_num_threads(void);
void orc_destroy(orc_t* o);
#ifdef __cplusplus
}
#endif
#endif
