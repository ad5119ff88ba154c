/*
 * wslda_td_project.h -- what a project library (a problem's predefines.h and
 * problem-definition.h compiled into libwslda_td.so by
 * `python -m wslda_amd.build --project DIR --codedim D`) exports beyond
 * wslda_td.h: the problem's own host functions, which the driver calls as
 * the reference's pca.c does (pca.c:876-914), and the run inputs they read.
 */
#ifndef WSLDA_TD_PROJECT_H
#define WSLDA_TD_PROJECT_H

#include <stddef.h>

#include "wslda_td.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The input-file values a problem's host code reads as input->name
 * (pca_utils.h:15-156, the fields problems use); the driver fills them
 * before calling process_params and reads sclgth back after it
 * (td-project-template/problem-definition.h:17-24). */
typedef struct {
    double sclgth, akF, dt, ec, kc, temperature, referencekF;
    int inittype, measurements, timesteps;
    double params[TDW_MAX_USER_PARAMS];
} tdw_project_inputs;

tdw_project_inputs *tdw_project_input(void);

/* The problem's functions (td-project-template/problem-definition.h). */
void process_params(double *params, double *kF, double *mu, size_t extra_data_size,
                    void *extra_data);                                  /* problem-definition.h:17 */
size_t get_extra_data_size(double *params);                             /* problem-definition.h:239 */
int load_extra_data(size_t extra_data_size, void *extra_data, double *params); /* problem-definition.h:253 */

#ifdef __cplusplus
}
#endif
#endif
