/*
 * sw2_internal.h -- declarations shared by the host C files and the HIP
 * device file of libsw2.  Not part of the public C-ABI.
 */
#ifndef SW2_INTERNAL_H
#define SW2_INTERNAL_H

#include "../../include/sw2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

void sw2_set_error(const char *fmt, ...);

/* Per-row tables uploaded to the device.  Each table has ny+4 entries,
 * index j+1 for rows j = -1 .. ny+2.  The geometry is j-only in channel
 * mode (coords.c:53-59, init.c:285-296), so a row table holds exactly the
 * value SW2 stores at every i of that row.  The transcendental entries are
 * computed on the host with the host libm, as SW2 computes them. */
enum {
    SW2_T_DXE = 0, SW2_T_DXC, SW2_T_DXV, SW2_T_DYN, SW2_T_DYC, SW2_T_DYV,
    SW2_T_GE, SW2_T_GN, SW2_T_GC, SW2_T_FE, SW2_T_FN,
    SW2_T_SINC, SW2_T_RZC, SW2_T_SINN, SW2_T_RZN, SW2_T_RMC, SW2_T_RMN,
    SW2_NT
};
/* Fill tab[SW2_NT*(ny+4)] from the world. */
void sw2_row_tables(sw *SW, double *tab);

/* Gravity-wave fallback of compute_Dt (calc.c:721-729): the two maxima over
 * the interior when the flow is at rest. */
void sw2_cfl_rest(sw *SW, double *max_x, double *max_y);

/* Host side of the in-loop perturbations.  For step time t, returns the
 * number of active storm injections and, for injection k, fills
 *   inc[k]   : eta increment per interior cell (S entries, 0 elsewhere),
 *   trval[k] : tracer value assigned where mask[k] != 0,
 *   mask[k]  : cells inside the truncation radius.
 * Mirrors add_Gaussians / add_Gaussian_core / get_Gaussian_field
 * (perturbs.c:22-395).  Returns -1 on SW2's fatal conditions. */
int sw2_gaussians_at(sw *SW, double Dt, double t, int maxk, double **inc, double **trval, unsigned char **mask);
/* Vortex increments at time t (perturbs.c:397-460): returns the number of
 * vortices starting in (t-Dt, t]; increments are summed per vortex k in
 * eta/u/v order into deta[k], du[k], dv[k]. */
int sw2_vortices_at(sw *SW, double Dt, double t, int maxk, double **deta, double **du, double **dv);

/* Check whether a field is identically zero (halos included). */
int sw2_all_zero(const double *f, int n);

/* Preprocess a case file (m4 subset); *out is malloc'ed. */
int sw2_preprocess(const char *fname, char **out, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
