/*
 * sphnj_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The FP64 CPU oracle: a from-scratch restatement of SPHNJ's per-step numerics (niji127/SPHNJ
 * src/{pair,sort,fluid,virt,solid,calculation}.cu) used to check libsphnj.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker
 * or the timed CPU baseline; the product path never links or calls it.
 *
 * The API mirrors include/sphnj.h (same parameter / particle structs, same stage numbers, same
 * download layout, same iteration orders), so a test drives both sides with the same calls.
 */
#ifndef SPHNJ_ORACLE_H
#define SPHNJ_ORACLE_H

#include "../include/sphnj.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

int oracle_create(const sphnj_params *per_case, int32_t ncase, oracle_ctx **out);
int oracle_upload(oracle_ctx *o, int32_t case_index, const sphnj_fluid_in *f, const sphnj_solid_in *s,
                  const sphnj_wall_in *w);
int oracle_init(oracle_ctx *o);
int oracle_step(oracle_ctx *o, int32_t nsteps);
int oracle_stage(oracle_ctx *o, int32_t stage);
int oracle_download(oracle_ctx *o, int32_t case_index, const sphnj_fluid_out *f, const sphnj_solid_out *s,
                    const sphnj_wall_out *w);
int oracle_status(oracle_ctx *o, uint32_t *flags);
int oracle_stage_sorted(oracle_ctx *o, int32_t case_index, int32_t species, int32_t *keys, int32_t *orig,
                        int32_t *cell_start);
int oracle_stage_neighbors(oracle_ctx *o, int32_t case_index, int32_t *counts, int32_t *list, int64_t cap,
                           int64_t *total);
int oracle_solid_reflist(oracle_ctx *o, int32_t case_index, int32_t *offsets, int32_t *list, int64_t cap);
int oracle_energy(oracle_ctx *o, int32_t case_index, double out[6]);
int oracle_time(oracle_ctx *o, int32_t case_index, double *time, double *last_dt);
int oracle_probe_global(oracle_ctx *o, int32_t case_index, double height_x, double out[SPHNJ_PROBE_COUNT]);
int oracle_probe_solid(oracle_ctx *o, int32_t case_index, double X, double Y, double pos[2], int32_t *index);
/* Derived constants of case 0 for the host-logic tests:
 * out = {rc2, hinv2, kdc, k2c, eR, nx, ny, cdx, cdy, lambda, mu_s, c_s, yield_strain, aux_modulus, umax_cap} */
int oracle_derived(oracle_ctx *o, double out[15]);
/* Number of OpenMP threads used (0 = runtime default); returns the count in effect. */
int oracle_set_threads(int32_t n);
const char *oracle_last_error(void);
void oracle_destroy(oracle_ctx *o);

#ifdef __cplusplus
}
#endif
#endif
