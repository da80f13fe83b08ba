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
    /* derived on the host, each with SW2's own expression and association:
     * reciprocals RN(1/x) of the divisors (the GPU divides through them,
     * exactly), 1.0/x factors SW2 writes as such (laplacian.c:58-89), the
     * row-constant quotients of laplacian_eta and do_eta_geofactor */
    SW2_T_IDXE, SW2_T_IDXC, SW2_T_IDXV, SW2_T_IDYN, SW2_T_IDYC, SW2_T_IDYV,
    SW2_T_IRZC, SW2_T_IRZN,
    SW2_T_DXE2, SW2_T_IDXE2, /* Dxe+Dxe (adv_Pu2, calc.c:198) and RN(1/.) */
    SW2_T_DYN2, SW2_T_IDYN2, /* Dyn(j)+Dyn(j+1) (adv_Pv2, calc.c:235) and RN(1/.) */
    SW2_T_FP,                /* rM(latc)/rZ(latc) (laplacian.c:121) */
    SW2_T_FX,                /* rZ(latn)/rM(latn) (laplacian.c:122) */
    SW2_T_KK,                /* 1.0/(rM(latc)*rZ(latc)) (laplacian.c:126) */
    SW2_T_GF,                /* (Dt*sin(latc))/rZ(latc) (calc.c:420) */
    SW2_NT
};
/* Fill tab[SW2_NT*(ny+4)] from the world.  Returns 1 when every divisor the
 * GPU divides through by reciprocal (rows 0 .. ny+1) lies in the range that
 * division accepts, 0 otherwise. */
int sw2_row_tables(sw *SW, double *tab);

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
/* The same injections as sparse lists: for injection k, the cells inside
 * its truncation radius in SW2's row-major order (inj[k].idx, field
 * offsets), the eta increment there (inj[k].inc; has_inc 0 when SW2 leaves
 * eta alone) and the tracer value (inj[k].trv).  Only those cells are
 * evaluated: per row the accepted columns are the ones whose cos(dlon) is
 * largest (the distance is monotone in it), found by walking the columns
 * in that order; every value, and both volume sums, are SW2's bit for bit.
 * The arrays grow as needed and stay owned by inj (sw2_inj_free). */
typedef struct {
    int n, cap, has_inc;
    int *idx;
    double *inc, *trv;
} sw2_inj;
int sw2_gaussians_sparse(sw *SW, double Dt, double t, int maxk, sw2_inj *inj);
void sw2_inj_free(sw2_inj *inj, int k);

/* Vortex increments at time t (perturbs.c:397-460): returns the number of
 * vortices starting in (t-Dt, t]; increments are summed per vortex k in
 * eta/u/v order into deta[k], du[k], dv[k]. */
int sw2_vortices_at(sw *SW, double Dt, double t, int maxk, double **deta, double **du, double **dv);

/* gf(j) = Dt*sin(latc)/rZ(latc) of do_eta_geofactor (calc.c:420) for the
 * rows j = -1..ny+2 (index j+1), computed with the host libm as SW2 does. */
void sw2_geofactor_rows(sw *SW, double Dt, double *gf);

/* Check whether a field is identically zero (halos included). */
int sw2_all_zero(const double *f, int n);

/* Preprocess a case file (m4 subset); *out is malloc'ed. */
int sw2_preprocess(const char *fname, char **out, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
