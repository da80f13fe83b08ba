/*
 * sw2_host.c -- host side of libsw2: the SW2 boundary (initialize / loop /
 * destroy), the sppde helpers SW2's driver links against, world set-up,
 * geometry, frame IO, host diagnostics and the host half of the in-loop
 * perturbations.  The per-step numerics run on the GPU (sw2_device.hip).
 *
 * Geometry and every host-side expression keep SW2's association so that
 * the values handed to the GPU are the values SW2 computes (sw.h:113-204,
 * coords.c:36-59, init.c:279-297).  Build flags: -O2 -ffp-contract=off, no
 * -march, so no multiply-add contraction happens on the host either.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <libgen.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "sw2_internal.h"

/* ------------------------------------------------------------------ errors */
static __thread char g_err[1024];

void sw2_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *sw2_last_error(void) { return g_err; }

/* ------------------------------------------- sppde helpers (single process) */
static FILE *g_out = NULL;
static int g_flushall = 0;
static char g_outpath[MAXOUTLEN];

int quants(void) { return 1; }
int quisoc(void) { return 0; }

void checkr(int r, char *s) {
    if (r != 0) crash(__LINE__, __FILE__, "checkr: %s returned %d", s, r);
}

void crash(int line, char *file, char *fmt, ...) {
    va_list ap;
    fflush(NULL);
    fprintf(stderr, "CRASH: proc=%d FILE=<%s> LINE=%d\n", 0, file, line);
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    fflush(NULL);
    exit(255); /* MPI_Abort(MPI_COMM_WORLD,-1) in sppde_core.c:183 */
}

void init_pprintf(int mode, int flushall, char *path) {
    (void)mode;
    g_flushall = flushall;
    struct stat st;
    if (path && stat(path, &st) == -1) {
        fprintf(stdout, "Processor 0 creating folder '%s' for output\n", path);
        char cmd[MAXOUTLEN + 16];
        snprintf(cmd, sizeof cmd, "%s", path);
        for (char *p = cmd + 1; *p; p++)
            if (*p == '/') {
                *p = '\0';
                mkdir(cmd, 0777);
                *p = '/';
            }
        if (mkdir(cmd, 0777) != 0 && errno != EEXIST) crash(__LINE__, __FILE__, "can't create output folder <%s>", path);
    }
    g_out = stdout;
    if (path) snprintf(g_outpath, sizeof g_outpath, "%s", path);
    if (flushall == 1) pprintf("Warning: fflush is called after each pprintf\n");
    pprintf("Warning: Matrix range is NOT checked\n");
}

void end_pprintf(void) {
    pprintf("End of output\n");
    fflush(NULL);
    g_out = NULL;
}

int pprintf(char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    int r = vfprintf(g_out ? g_out : stdout, fmt, ap);
    va_end(ap);
    if (g_flushall) fflush(g_out ? g_out : stdout);
    return r;
}

double *dmem(map *M) {
    double *r = (double *)malloc(sizeof(double) * (size_t)M->S);
    if (!r) crash(__LINE__, __FILE__, "oops out of memory S=%d", M->S);
    return r;
}

/* wall-clock channels (sppde_cr.c:38-207) */
#define NCH 80
static char ch_name[NCH][64];
static double ch_t0[NCH], ch_sum[NCH], ch_min[NCH], ch_max[NCH], ch_last[NCH];
static int ch_n[NCH], ch_on[NCH], nch = 0;

static double wtime(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
static int ch_find(const char *s, int level) {
    char key[64];
    snprintf(key, sizeof key, "%s%02d", s, level);
    for (int k = 0; k < nch; k++)
        if (strcmp(ch_name[k], key) == 0) return k;
    if (nch >= NCH) return -1;
    snprintf(ch_name[nch], sizeof ch_name[nch], "%s", key);
    ch_min[nch] = 1e300;
    return nch++;
}
void cr_reset(void) { nch = 0; }
void cr_start(char *s, int level) {
    int k = ch_find(s, level);
    if (k < 0) return;
    ch_on[k] = 1;
    ch_t0[k] = wtime();
}
void cr_end(char *s, int level) {
    int k = ch_find(s, level);
    if (k < 0 || !ch_on[k]) return;
    double d = wtime() - ch_t0[k];
    ch_on[k] = 0;
    ch_sum[k] += d;
    ch_last[k] = d;
    ch_n[k]++;
    if (d < ch_min[k]) ch_min[k] = d;
    if (d > ch_max[k]) ch_max[k] = d;
}
double cr_time(char *s, int level) {
    int k = ch_find(s, level);
    if (k < 0) return 0;
    return ch_on[k] ? wtime() - ch_t0[k] : ch_last[k];
}
void cr_info(void) {
    int used[NCH] = {0};
    pprintf("cr_info\tnom\to\tN\tmode\ttmin\ttmax\ttsum\ttavg(s)\n");
    for (int q = 0; q < nch; q++) {
        int pj = 0;
        double mx = -1.0;
        for (int k = 0; k < nch; k++)
            if (!used[k] && ch_sum[k] > mx) {
                mx = ch_sum[k];
                pj = k;
            }
        used[pj] = 1;
        pprintf("cr_info:\t%13s\t%2d\t%5d\t%2d\t%4d\t%e\t%e\t%e\t%e\n", ch_name[pj], pj, ch_on[pj] != 0, ch_n[pj], 0,
                ch_min[pj], ch_max[pj], ch_sum[pj], ch_sum[pj] / (1. * (ch_n[pj] ? ch_n[pj] : 1)));
    }
}

/* --------------------------------------------------------------- geometry */
#define PI_ (4.0 * atan(1.0))
#define DEG2RAD(a) ((a) * (PI_ / 180.0))
#define RAD2DEG(b) ((b) * (180.0 / PI_))

static double linspace_val(int i, double x0, double x1, int n) { return (x0 + (double)i * (x1 - x0) / ((double)n)); }
static double g_latn(sw *SW, int j) { return linspace_val(j, SW->lat0, SW->lat1, SW->ny); }
static double g_latc(sw *SW, int j) { return ((g_latn(SW, j) + g_latn(SW, j - 1)) / 2.); }
static double g_lone(sw *SW, int i) { return linspace_val(i, SW->lon0, SW->lon1, SW->nx); }
static double g_lonc(sw *SW, int i) { return ((g_lone(SW, i) + g_lone(SW, i - 1)) / 2.); }
static double g_rZ(sw *SW, double lat) { return ((SW->rE) / sqrt(1.0 + (tan(lat) * tan(lat)) / (SW->epsilon2))); }
static double g_rM(sw *SW, double lat) {
    double x = g_rZ(SW, lat) / ((SW->rE) * cos(lat));
    return ((SW->rE) / (SW->epsilon2) * (x * x * x));
}
static double g_lat2pclat(sw *SW, double lat) { return atan(1.0 / (SW->epsilon2) * tan(lat)); }
static double g_r(sw *SW, double p) {
    double a = SW->rE * sin(p), b = SW->rP * cos(p);
    return ((SW->rE) * (SW->rP) / sqrt(a * a + b * b));
}
static double g_g0(sw *SW, double p) {
    double q = (SW->rE) / (g_r(SW, p));
    return ((SW->gE + ((SW->Omega) * (SW->Omega)) * (SW->rE)) * (q * q));
}
static double g_geff(sw *SW, double p) { return (g_g0(SW, p) - ((SW->Omega) * (SW->Omega)) * g_r(SW, p) * cos(p)); }
static double g_f(sw *SW, double lat) { return (2.0 * (SW->Omega) * sin(lat)); }

static void fill_rows(sw *SW, double *field, double (*fn)(sw *, int)) {
    map *M = SW->M;
    for (int j = -1; j <= SW->ny + 2; j++) {
        double v = fn(SW, j);
        for (int i = -1; i <= SW->nx + 2; i++) ac(field, i, j, M) = v;
    }
}
static double r_ge(sw *SW, int j) { return g_geff(SW, g_lat2pclat(SW, g_latc(SW, j))); }
static double r_gn(sw *SW, int j) { return g_geff(SW, g_lat2pclat(SW, g_latn(SW, j))); }
static double r_fe(sw *SW, int j) { return g_f(SW, g_latc(SW, j)); }
static double r_fn(sw *SW, int j) { return g_f(SW, g_latn(SW, j)); }

/* init_coords_spheroid (coords.c:21-126): distances and coordinates */
static void coords(sw *SW) {
    map *M = SW->M;
    double dlon = (SW->lon1 - SW->lon0) / ((double)SW->nx);
    double dlat = (SW->lat1 - SW->lat0) / ((double)SW->ny);
    double yn = 0, yc = 0, yv = 0, deltax = 0, deltay = 0;
    for (int j = -1; j <= SW->ny + 2; j++) {
        double xe = 0, xc = 0, xv = 0;
        double dxe = dlon * g_rZ(SW, g_latc(SW, j));
        double dxc = dlon * g_rZ(SW, g_latc(SW, j));
        double dxv = dlon * g_rZ(SW, g_latn(SW, j));
        double dyn = dlat * g_rM(SW, g_latc(SW, j));
        double dyc = dlat * g_rM(SW, g_latn(SW, j - 1));
        double dyv = dlat * g_rM(SW, g_latc(SW, j));
        yn += dyn;
        yc += dyc;
        yv += dyv;
        for (int i = -1; i <= SW->nx + 2; i++) {
            xe += dxe;
            xc += dxc;
            xv += dxv;
            if (i == 0 && j == 0) {
                deltax = xe;
                deltay = yn;
            }
            ac(SW->Dxe, i, j, M) = dxe;
            ac(SW->Dxc, i, j, M) = dxc;
            ac(SW->Dxv, i, j, M) = dxv;
            ac(SW->Dyn, i, j, M) = dyn;
            ac(SW->Dyc, i, j, M) = dyc;
            ac(SW->Dyv, i, j, M) = dyv;
            ac(SW->xe, i, j, M) = xe;
            ac(SW->xc, i, j, M) = xc;
            ac(SW->xv, i, j, M) = xv;
            ac(SW->yn, i, j, M) = yn;
            ac(SW->yc, i, j, M) = yc;
            ac(SW->yv, i, j, M) = yv;
        }
    }
    for (int k = 0; k < M->S; k++) {
        SW->xe[k] -= deltax;
        SW->xc[k] -= deltax;
        SW->xv[k] -= deltax;
        SW->yn[k] -= deltay;
        SW->yc[k] -= deltay;
        SW->yv[k] -= deltay;
    }
}

void sw2_geofactor_rows(sw *SW, double Dt, double *gf) {
    for (int j = -1; j <= SW->ny + 2; j++) {
        double latc = g_latc(SW, j);
        gf[j + 1] = Dt * sin(latc) / g_rZ(SW, latc);
    }
}

int sw2_row_tables(sw *SW, double *tab) {
    int R = SW->ny + 4;
    map *M = SW->M;
    for (int j = -1; j <= SW->ny + 2; j++) {
        int r = j + 1;
        double latc = g_latc(SW, j), latn = g_latn(SW, j);
        tab[SW2_T_DXE * R + r] = ac(SW->Dxe, 1, j, M);
        tab[SW2_T_DXC * R + r] = ac(SW->Dxc, 1, j, M);
        tab[SW2_T_DXV * R + r] = ac(SW->Dxv, 1, j, M);
        tab[SW2_T_DYN * R + r] = ac(SW->Dyn, 1, j, M);
        tab[SW2_T_DYC * R + r] = ac(SW->Dyc, 1, j, M);
        tab[SW2_T_DYV * R + r] = ac(SW->Dyv, 1, j, M);
        tab[SW2_T_GE * R + r] = ac(SW->ge, 1, j, M);
        tab[SW2_T_GN * R + r] = ac(SW->gn, 1, j, M);
        tab[SW2_T_GC * R + r] = ac(SW->gc, 1, j, M);
        tab[SW2_T_FE * R + r] = ac(SW->fe, 1, j, M);
        tab[SW2_T_FN * R + r] = ac(SW->fn, 1, j, M);
        tab[SW2_T_SINC * R + r] = sin(latc);
        tab[SW2_T_RZC * R + r] = g_rZ(SW, latc);
        tab[SW2_T_SINN * R + r] = sin(latn);
        tab[SW2_T_RZN * R + r] = g_rZ(SW, latn);
        tab[SW2_T_RMC * R + r] = g_rM(SW, latc);
        tab[SW2_T_RMN * R + r] = g_rM(SW, latn);
    }
    static const int inv[][2] = {{SW2_T_IDXE, SW2_T_DXE}, {SW2_T_IDXC, SW2_T_DXC}, {SW2_T_IDXV, SW2_T_DXV},
                                 {SW2_T_IDYN, SW2_T_DYN}, {SW2_T_IDYC, SW2_T_DYC}, {SW2_T_IDYV, SW2_T_DYV},
                                 {SW2_T_IRZC, SW2_T_RZC}, {SW2_T_IRZN, SW2_T_RZN}};
    for (int r = 0; r < R; r++) {
        for (unsigned k = 0; k < sizeof inv / sizeof *inv; k++) tab[inv[k][0] * R + r] = 1.0 / tab[inv[k][1] * R + r];
        double dxe = tab[SW2_T_DXE * R + r];
        tab[SW2_T_DXE2 * R + r] = dxe + dxe;
        tab[SW2_T_IDXE2 * R + r] = 1.0 / (dxe + dxe);
        double dyn2 = (r + 1 < R) ? tab[SW2_T_DYN * R + r] + tab[SW2_T_DYN * R + r + 1] : 1.0;
        tab[SW2_T_DYN2 * R + r] = dyn2;
        tab[SW2_T_IDYN2 * R + r] = 1.0 / dyn2;
        double rmc = tab[SW2_T_RMC * R + r], rzc = tab[SW2_T_RZC * R + r];
        tab[SW2_T_FP * R + r] = rmc / rzc;
        tab[SW2_T_FX * R + r] = tab[SW2_T_RZN * R + r] / tab[SW2_T_RMN * R + r];
        tab[SW2_T_KK * R + r] = 1.0 / (rmc * rzc);
        tab[SW2_T_GF * R + r] = SW->Dt * tab[SW2_T_SINC * R + r] / rzc;
    }
    /* divisors taken through their reciprocals: positive, in [2^-60, 2^60] */
    static const int dv[] = {SW2_T_DXE, SW2_T_DXC, SW2_T_DXV, SW2_T_DYN, SW2_T_DYC, SW2_T_DYV,
                             SW2_T_RZC, SW2_T_RZN, SW2_T_DXE2, SW2_T_DYN2};
    int ok = 1;
    for (int r = 1; r <= SW->ny + 2; r++) /* rows j = 0 .. ny+1 */
        for (unsigned k = 0; k < sizeof dv / sizeof *dv; k++) {
            double b = tab[dv[k] * R + r];
            if (!(b >= 0x1p-60 && b <= 0x1p+60)) ok = 0;
        }
    return ok;
}

void sw2_cfl_rest(sw *SW, double *mx, double *my) {
    double max_x = 0, max_y = 0;
    map *M = SW->M;
    for (int j = 1; j <= SW->ny; j++)
        for (int i = 1; i <= SW->nx; i++) {
            if (fabs(sqrt((SW->depth) * (ac(SW->ge, i, j, M))) / ac(SW->Dxe, i, j, M)) > max_x)
                max_x = fabs(sqrt((SW->depth) * (ac(SW->ge, i, j, M)))) / ac(SW->Dxe, i, j, M);
            if (fabs(sqrt((SW->depth) * (ac(SW->gn, i, j, M))) / ac(SW->Dyn, i, j, M)) > max_y)
                max_y = fabs(sqrt((SW->depth) * (ac(SW->gn, i, j, M)))) / ac(SW->Dyn, i, j, M);
        }
    *mx = max_x;
    *my = max_y;
}

/* ------------------------------------------- halo rules, channel, energy */
void halo_update(double *x, map *M) {
    int nx = M->l1[1], ny = M->l1[2];
    for (int i = 1; i <= nx; i++) { /* sppde_hupdate.c:296-304 */
        ac(x, i, -1, M) = ac(x, i, ny - 1, M);
        ac(x, i, 0, M) = ac(x, i, ny, M);
        ac(x, i, ny + 1, M) = ac(x, i, 1, M);
        ac(x, i, ny + 2, M) = ac(x, i, 2, M);
    }
    for (int j = -1; j <= ny + 2; j++) { /* sppde_hupdate.c:335-339 */
        ac(x, -1, j, M) = ac(x, nx - 1, j, M);
        ac(x, 0, j, M) = ac(x, nx, j, M);
        ac(x, nx + 1, j, M) = ac(x, 1, j, M);
        ac(x, nx + 2, j, M) = ac(x, 2, j, M);
    }
}

void do_channel(sw *SW) {
    map *M = SW->M;
    int ny = SW->ny;
    for (int i = 1; i <= SW->nx; i++) { /* calc.c:614-629 */
        ac(SW->v, i, 0, M) = 0;
        ac(SW->v, i, -1, M) = 0;
        ac(SW->windy, i, 0, M) = 0;
        ac(SW->windy, i, -1, M) = 0;
        ac(SW->u, i, 0, M) = ac(SW->u, i, 1, M);
        ac(SW->u, i, -1, M) = ac(SW->u, i, 0, M);
        ac(SW->windx, i, 0, M) = ac(SW->windx, i, 1, M);
        ac(SW->windx, i, -1, M) = ac(SW->windx, i, 0, M);
    }
    for (int i = 1; i <= SW->nx; i++) { /* calc.c:631-647 */
        ac(SW->v, i, ny, M) = 0;
        ac(SW->v, i, ny + 1, M) = 0;
        ac(SW->v, i, ny + 2, M) = 0;
        ac(SW->windy, i, ny, M) = 0;
        ac(SW->windy, i, ny + 1, M) = 0;
        ac(SW->windy, i, ny + 2, M) = 0;
        ac(SW->u, i, ny + 1, M) = ac(SW->u, i, ny, M);
        ac(SW->u, i, ny + 2, M) = ac(SW->u, i, ny + 1, M);
        ac(SW->windx, i, ny + 1, M) = ac(SW->windx, i, ny, M);
        ac(SW->windx, i, ny + 2, M) = ac(SW->windx, i, ny + 1, M);
    }
}

#define HH(i, j) ((ac(SW->eta, i, j, M)) + (SW->depth) - (ac(SW->hB, i, j, M)))
#define UW(i, j) (ac(SW->u, i, j, M) + ac(SW->windx, i, j, M))
#define VW(i, j) (ac(SW->v, i, j, M) + ac(SW->windy, i, j, M))

void calc_energy_balance(double *ke, double *ape, sw *SW) { /* energy.c:25-55 */
    map *M = SW->M;
    double myke = 0, myape = 0, uc, vc, mysurf;
    for (int j = 1; j <= SW->ny; j++)
        for (int i = 1; i <= SW->nx; i++) {
            uc = 0.5 * (UW(i, j) + UW(i - 1, j));
            vc = 0.5 * (VW(i, j) + VW(i, j - 1));
            mysurf = ac(SW->Dxc, i, j, M) * ac(SW->Dyc, i, j, M);
            double wx = ac(SW->windx, i, j, M), wy = ac(SW->windy, i, j, M);
            myke += 0.5 * HH(i, j) * (uc * uc + vc * vc) * mysurf - 0.5 * ((wx * wx) + (wy * wy)) * mysurf;
            myape += 0.5 * ac(SW->gc, i, j, M) * (HH(i, j) * HH(i, j)) * mysurf - 0.5 * ac(SW->gc, i, j, M) * ((SW->depth) * (SW->depth)) * mysurf;
        }
    *ke = myke;
    *ape = myape;
}

double sw2_mass(sw *SW) {
    map *M = SW->M;
    double m = 0;
    for (int j = 1; j <= SW->ny; j++)
        for (int i = 1; i <= SW->nx; i++) m += HH(i, j) * (ac(SW->Dxc, i, j, M) * ac(SW->Dyc, i, j, M));
    return m;
}

static void print_energy(double t, double ke, double ape) { pprintf("ENERGY %e %e %e %e\n", t, ke, ape, ke + ape); }

static void get_stats(sw *SW, double *r, double *gmin, double *gmax, double *gavg) { /* sppde_core.c:346-375 */
    map *M = SW->M;
    double s = 0, n = 0, mn = 0, mx = 0;
    for (int j = 1; j <= SW->ny; j++)
        for (int i = 1; i <= SW->nx; i++) {
            n = n + 1;
            s += ac(r, i, j, M);
            if (ac(r, i, j, M) < mn || n == 1.0) mn = ac(r, i, j, M);
            if (ac(r, i, j, M) > mx || n == 1.0) mx = ac(r, i, j, M);
        }
    *gmin = mn;
    *gmax = mx;
    *gavg = s / n;
}

static void stats_line(sw *SW, double *r, const char *name) {
    double a, b, c;
    get_stats(SW, r, &a, &b, &c);
    pprintf("stats_scaf fname=<%s> min=%e max=%e avg=%e \n", name, a, b, c);
}

int sw2_all_zero(const double *f, int n) {
    for (int k = 0; k < n; k++)
        if (f[k] != 0.0) return 0;
    return 1;
}

/* ------------------------------------------------------------- world set-up */
/* createM for one process (sppde_core.c:461-594), with the values it gives
 * there, except the extent of the third axis: SW2 sizes it with five planes
 * of halo that no 2-D loop reads (s[3] = 6, so every field is six times the
 * 2-D array); here s[3] = 1 and S is the 2-D field, which is all that
 * LP / ac and every loop over (i, j) address. */
static void single_map(sw *SW) {
    map *M = &SW->m;
    memset(M, 0, sizeof *M);
    M->sh = 2;
    M->nd = 2;
    M->l0[1] = M->gl0[1] = 1;
    M->l1[1] = M->gl1[1] = SW->nx;
    M->l0[2] = M->gl0[2] = 1;
    M->l1[2] = M->gl1[2] = SW->ny;
    M->gl0[3] = M->gl1[3] = 1;
    M->l0[3] = 0;
    M->l1[3] = 1;
    M->al0[0][1] = 1, M->al1[0][1] = SW->nx;
    M->al0[0][2] = 1, M->al1[0][2] = SW->ny;
    M->al0[0][3] = 0, M->al1[0][3] = 1;
    M->pco[0][1] = M->pco[0][2] = M->pco[0][3] = 1;
    M->s[1] = SW->nx + 4;
    M->s[2] = SW->ny + 4;
    M->s[3] = 1;
    M->S = M->s[1] * M->s[2];
    M->gpc[1] = M->gpc[2] = M->gpc[3] = 1;
    M->pc[1] = M->pc[2] = M->pc[3] = 1;
    M->nb_e = M->nb_w = 0; /* periodic x: own neighbour */
    M->nb_n = M->nb_s = -1;
    M->nb_t = M->nb_b = -1;
    M->per[1] = 1;
    M->per[2] = 0;
    M->bc0[1] = M->bc1[1] = 0; /* periodic axis owns no boundary */
    M->bc0[2] = M->bc1[2] = 1;
    M->bc0[3] = M->bc1[3] = 1;
    SW->M = M;
}

static double *zfield(sw *SW) {
    double *f = (double *)calloc((size_t)SW->M->S, sizeof(double));
    if (!f) sw2_set_error("out of memory for a %d-double field", SW->M->S);
    return f;
}

static int linterp_host(double xi, const double *x, const double *y, int n, double *out) { /* tools.c:21-60 */
    int i = 0, i1 = n - 1;
    if (x[0] > xi || x[i1] < xi) return -1;
    while ((i < i1) && (x[i + 1] <= xi) && (i >= 0)) i++;
    if (i == i1) i = i1 - 1; /* SW2 would read past the table here */
    *out = (y[i + 1] - y[i]) * (xi - x[i]) / (x[i + 1] - x[i]) + y[i];
    return 0;
}

int sw2_init_params(const sw2_params *P, sw *SW) {
    memset(SW, 0, sizeof *SW);
    if (P->polar != 0) {
        sw2_set_error("polar=%d: only the channel geometry runs on the GPU path", P->polar);
        return -1;
    }
    if (P->nx < 3 || P->ny < 3) {
        sw2_set_error("grid %dx%d too small (at least 3 cells per axis)", P->nx, P->ny);
        return -1;
    }
    if (P->sch < 0 || P->sch > 9) {
        sw2_set_error("scheme=%d a la parra", P->sch);
        return -1;
    }
    SW->sch = P->sch;
    SW->polar = 0;
    SW->lon0 = DEG2RAD(P->lon0_deg); /* init.c:56-59 */
    SW->lon1 = DEG2RAD(P->lon1_deg);
    SW->lat0 = DEG2RAD(P->lat0_deg);
    SW->lat1 = DEG2RAD(P->lat1_deg);
    SW->nx = P->nx;
    SW->ny = P->ny;
    SW->nz = 1;
    single_map(SW);
    SW->t0 = P->t0;
    SW->t1 = P->t1;
    SW->Dt = P->Dt;
    SW->Courant_max = P->CFL;
    SW->N = (int)(((SW->t1) - (SW->t0)) / (SW->Dt)); /* init.c:109 */
    SW->sigma_dis = (P->Tau > 1e10) ? 0 : 1.0 / P->Tau;
    SW->nu2 = P->hypernu[0];
    SW->nu4 = -P->hypernu[1];
    SW->nu6 = P->hypernu[2];
    if (SW->nu2 < 0 || SW->nu4 > 0 || SW->nu6 < 0) {
        sw2_set_error("Hyperviscosity coefs should be entered as >0 and they are %e %e %e", SW->nu2, -SW->nu4, SW->nu6);
        return -1;
    }
    SW->IteInfo = P->IteInfo > 0 ? P->IteInfo : 1;
    SW->saveEvery = P->SaveEvery > 0 ? P->SaveEvery : 1;
    SW->loadFrom = P->LoadFrom;
    SW->reftype = P->reftype;
    SW->Omega = P->Omega;
    SW->rE = P->rE;
    SW->rP = P->rP;
    SW->epsilon2 = ((SW->rE) / (SW->rP)) * ((SW->rE) / (SW->rP));
    SW->gE = P->gE;
    SW->depth = P->depth;
    SW->beta_sponge = 1e5;
    SW->sigma_sponge = 0;
    SW->tau_sponge = 0;
    SW->geoeq = P->geoeq;
    SW->tracer_dis = P->TracerDis;
    double **all[] = {&SW->Dxe, &SW->Dxc, &SW->Dxv, &SW->Dyn, &SW->Dyc, &SW->Dyv, &SW->xe, &SW->xc, &SW->xv, &SW->yn,
                      &SW->yc, &SW->yv, &SW->ge, &SW->gn, &SW->gc, &SW->gv, &SW->fe, &SW->fn, &SW->fc, &SW->fv,
                      &SW->tracer, &SW->perturbs, &SW->u, &SW->v, &SW->windx, &SW->windy, &SW->eta, &SW->hB};
    for (size_t k = 0; k < sizeof all / sizeof all[0]; k++)
        if (!(*all[k] = zfield(SW))) return -1;
    fill_rows(SW, SW->ge, r_ge); /* init_channel_region, init.c:279-297 */
    fill_rows(SW, SW->gn, r_gn);
    fill_rows(SW, SW->gc, r_ge);
    fill_rows(SW, SW->gv, r_gn);
    fill_rows(SW, SW->fe, r_fe);
    fill_rows(SW, SW->fn, r_fn);
    fill_rows(SW, SW->fc, r_fe);
    fill_rows(SW, SW->fv, r_fn);
    coords(SW);
    SW->num_vortices = P->num_vortices;
    SW->vortices = (double *)calloc((size_t)P->num_vortices * N_COLS_VORTICES + 1, sizeof(double));
    if (P->num_vortices) memcpy(SW->vortices, P->vortices, sizeof(double) * P->num_vortices * N_COLS_VORTICES);
    SW->num_Gaussians = P->num_Gaussians;
    SW->Gaussians = (double *)calloc((size_t)P->num_Gaussians * N_COLS_GAUSS + 1, sizeof(double));
    if (P->num_Gaussians) memcpy(SW->Gaussians, P->Gaussians, sizeof(double) * P->num_Gaussians * N_COLS_GAUSS);
    if (P->wind_rows > 0) { /* read_wind, io.c:173-205 */
        int nr = P->wind_rows;
        double *x = (double *)malloc(sizeof(double) * nr), *y = (double *)malloc(sizeof(double) * nr);
        for (int r = 0; r < nr; r++) {
            x[r] = P->wind[2 * r];
            y[r] = P->wind[2 * r + 1];
        }
        /* the interpolant depends on the row only: one table search per row */
        for (int j = -1; j <= SW->ny + 2; j++) {
            double w;
            if (linterp_host(RAD2DEG(g_latc(SW, j)), x, y, nr, &w)) {
                sw2_set_error("Latitude value %e out of bounds of the wind table (%e, %e)", RAD2DEG(g_latc(SW, j)), x[0], x[nr - 1]);
                free(x);
                free(y);
                return -1;
            }
            for (int i = -1; i <= SW->nx + 2; i++) ac(SW->windx, i, j, SW->M) = w;
        }
        free(x);
        free(y);
    }
    halo_update(SW->windx, SW->M); /* init.c:232-235 */
    halo_update(SW->windy, SW->M);
    return 0;
}

int initialize(int npx, int npy, int force1proc, char *bname, int copyinput, sw *SW) {
    char fname[MAXOUTLEN + 8];
    snprintf(fname, sizeof fname, "%s.sw", bname);
    if (copyinput == 1 && g_outpath[0]) { /* sppde_parser.c:185-197: copy, then the expanded text */
        char *txt = NULL, err[256];
        FILE *fi = fopen(fname, "rb");
        if (fi) {
            char dst[2 * MAXOUTLEN + 16];
            snprintf(dst, sizeof dst, "%s/%s", g_outpath, basename(fname));
            FILE *fo = fopen(dst, "wb");
            if (fo) {
                char buf[4096];
                size_t k;
                while ((k = fread(buf, 1, sizeof buf, fi)) > 0) fwrite(buf, 1, k, fo);
                fclose(fo);
            }
            fclose(fi);
        }
        if (sw2_preprocess(fname, &txt, err, sizeof err) == 0) {
            char dst[2 * MAXOUTLEN + 16];
            snprintf(dst, sizeof dst, "%s/%s.prem4", g_outpath, basename(fname));
            FILE *f = fopen(dst, "w");
            if (f) {
                fputs(txt, f);
                fclose(f);
            }
            free(txt);
        }
    }
    sw2_params P;
    if (sw2_read_case(fname, &P)) CRASH("%s", sw2_last_error());
    pprintf("initialize npx= %d npy= %d kN= %.1f kN/P= %.1f\n", npx, npy, 1e-3 * P.nx * P.ny, 1e-3 * P.nx * P.ny / (npx * npy));
    if (force1proc == 1) {
        npx = 1;
        npy = 1;
    }
    if (npx * npy != 1)
        CRASH("oops quants=%d doesnt agree with npx npy npz = %d %d %d. Run with %d procs (one GPU owns the whole domain: use npx=npy=1)",
              quants(), npx, npy, 1, npx * npy);
    if (sw2_init_params(&P, SW)) CRASH("%s", sw2_last_error());
    if (SW->reftype != 0) CRASH("reftype=%d: comparison against reference text frames is not supported by this build", SW->reftype);
    pprintf("Gaussians table read. There is/are %d gaussian perturbations in the domain.\n", SW->num_Gaussians);
    pprintf("Zonal winds added.\n");
    sw2_free_params(&P);
    return 0;
}

void destroy(sw *SW) {
    double *all[] = {SW->Dxe, SW->Dxc, SW->Dxv, SW->Dyn, SW->Dyc, SW->Dyv, SW->xe, SW->xc, SW->xv, SW->yn,
                     SW->yc, SW->yv, SW->ge, SW->gn, SW->gc, SW->gv, SW->fe, SW->fn, SW->fc, SW->fv,
                     SW->tracer, SW->perturbs, SW->u, SW->v, SW->windx, SW->windy, SW->eta, SW->hB,
                     SW->Gaussians, SW->vortices, SW->reftable};
    for (size_t k = 0; k < sizeof all / sizeof all[0]; k++) free(all[k]);
    memset(SW, 0, sizeof *SW);
}

/* ------------------------------------------------------------- frame IO */
#define MAGICSW1 27
#define MAGICSW2 28
#define MAGIC_SPPDE 15

static void fwrite_field(double *x, FILE *f, sw *SW) { /* sppde_binary_files.c:36-60 */
    int magic = MAGIC_SPPDE;
    if (fwrite(&magic, sizeof(int), 1, f) != 1) CRASH("oops problems writing binary file magic number");
    if (fwrite(x, sizeof(double), (size_t)SW->M->S, f) != (size_t)SW->M->S) CRASH("oops problems writing binary file");
}

void save_binary_sw(int frame, int nstep, double timedouble, char *bname, char *output_folder, sw *SW) {
    char outname[2 * MAXOUTLEN + 16];
    snprintf(outname, sizeof outname, "%s/%s.%04d.bin", output_folder, basename(bname), frame);
    FILE *fout = fopen(outname, "w");
    if (!fout) CRASH("file cannot be opened <%s>", outname);
    int magic = MAGICSW2;
    fwrite(&magic, sizeof(int), 1, fout);
    fwrite(&nstep, sizeof(int), 1, fout);
    fwrite(&timedouble, sizeof(double), 1, fout);
    fwrite(&SW->nx, sizeof(int), 1, fout);
    fwrite(&SW->ny, sizeof(int), 1, fout);
    pprintf("save_binary_sw saved frame=%d nstep=%d timedouble=%e outname=<%s> \n", frame, nstep, timedouble, outname);
    fwrite_field(SW->eta, fout, SW);
    fwrite_field(SW->u, fout, SW);
    fwrite_field(SW->v, fout, SW);
    fwrite_field(SW->tracer, fout, SW);
    fwrite_field(SW->perturbs, fout, SW);
    fclose(fout);
}

static void fread_field(double *x, FILE *f, sw *SW) { /* sppde_binary_files.c:147-180 */
    int magic = 0;
    if (fread(&magic, sizeof(int), 1, f) != 1) CRASH("oops problems reading binary file magic number");
    if (magic != MAGIC_SPPDE) CRASH("Wrong binary field magic, is %d and should be %d", magic, MAGIC_SPPDE);
    double *g = (double *)malloc(sizeof(double) * (size_t)SW->M->S);
    if (fread(g, sizeof(double), (size_t)SW->M->S, f) != (size_t)SW->M->S) CRASH("oops problems reading binary file");
    map *M = SW->M;
    for (int j = 1; j <= SW->ny; j++) /* scatter interior only, then halo update */
        for (int i = 1; i <= SW->nx; i++) ac(x, i, j, M) = ac(g, i, j, M);
    free(g);
    halo_update(x, M);
}

int load_binary_sw(int frame, int *nstep, double *timedouble, char *bname, char *binaries_folder, sw *SW) {
    char inname[2 * MAXOUTLEN + 16];
    snprintf(inname, sizeof inname, "%s/%s.%04d.bin", binaries_folder, basename(bname), frame);
    FILE *fin = fopen(inname, "r");
    if (!fin) return 1;
    int magic, nxr, nyr;
    if (fread(&magic, sizeof(int), 1, fin) != 1) CRASH("File to be loaded is corrupted");
    if (magic < MAGICSW1 || magic > MAGICSW2) CRASH("Wrong magic = %d should be %d or %d", magic, MAGICSW1, MAGICSW2);
    if (fread(nstep, sizeof(int), 1, fin) != 1) CRASH("File to be loaded is corrupted");
    if (fread(timedouble, sizeof(double), 1, fin) != 1) CRASH("File to be loaded is corrupted");
    if (fread(&nxr, sizeof(int), 1, fin) != 1) CRASH("File to be loaded is corrupted");
    if (fread(&nyr, sizeof(int), 1, fin) != 1) CRASH("File to be loaded is corrupted");
    if (nxr != SW->nx) CRASH("oops nxr=%d should be %d", nxr, SW->nx);
    if (nyr != SW->ny) CRASH("oops nyr=%d should be %d", nyr, SW->ny);
    if (magic == MAGICSW1) {
        fread_field(SW->eta, fin, SW);
        fread_field(SW->u, fin, SW);
        fread_field(SW->v, fin, SW);
        fread_field(SW->perturbs, fin, SW);
        memset(SW->tracer, 0, sizeof(double) * (size_t)SW->M->S);
    } else {
        fread_field(SW->eta, fin, SW);
        fread_field(SW->u, fin, SW);
        fread_field(SW->v, fin, SW);
        fread_field(SW->tracer, fin, SW);
        fread_field(SW->perturbs, fin, SW);
    }
    fclose(fin);
    return 0;
}

/* ------------------------------------------------- in-loop perturbations */
static double great_circle(double lon0, double lat0, double lon1, double lat1, double r) {
    return (r) * acos(sin(lat0) * sin(lat1) + cos(lat0) * cos(lat1) * cos(lon1 - lon0));
}

/* get_Gaussian_field (perturbs.c:348-395) on host scratch arrays */
static void gaussian_field(sw *SW, double lon, double lat, double amp, double sigma, double radius, double *field,
                           double inject, double *trval, unsigned char *mask) {
    map *M = SW->M;
    if (SW->prn) pprintf("get_Gaussian_field entering lon=%f DEG lat=%f DEG\n", RAD2DEG(lon), RAD2DEG(lat));
    for (int j = 1; j <= SW->ny; j++)
        for (int i = 1; i <= SW->nx; i++) {
            double r = great_circle(lon, g_lat2pclat(SW, lat), g_lonc(SW, i), g_lat2pclat(SW, g_latc(SW, j)), (SW->rE + SW->rP) / 2.0);
            if (r < radius) {
                ac(field, i, j, M) += amp * exp(-(r * r) / (2.0 * (sigma * sigma)));
                ac(trval, i, j, M) = (inject) * exp(-(r * r) / (2.0 * (sigma * sigma)));
                mask[LP(i, j, M)] = 1;
            }
        }
}

int sw2_gaussians_at(sw *SW, double Dt, double t, int maxk, double **inc, double **trval, unsigned char **mask) {
    int K = 0;
    map *M = SW->M;
    for (int j = 0; j <= SW->num_Gaussians - 1; j++) {
        double *g = SW->Gaussians + j * N_COLS_GAUSS;
        double t0g = g[0], t1g = g[1];
        if (!((t0g <= t) && (t1g >= t))) continue;
        if (K >= maxk) {
            sw2_set_error("more than %d simultaneous injections", maxk);
            return -1;
        }
        double tau = g[2];
        if (tau < 0) {
            sw2_set_error("Uhhh tau is %e and should be positive", tau);
            return -1;
        }
        double lon_g = DEG2RAD(g[3]), lat_g = DEG2RAD(g[4]);
        int isdegvel = (int)g[5];
        double vel = g[6];
        if (isdegvel < 0 || isdegvel > 1) {
            sw2_set_error("Perturbation table error: isdegvel should be 0 or 1");
            return -1;
        }
        double tt = (t - t0g), Om = 0;
        if (isdegvel == 0) Om = vel / g_rZ(SW, lat_g);
        if (isdegvel == 1) Om = DEG2RAD(vel);
        double lon_g0 = lon_g;
        lon_g = lon_g0 + Om * tt;
        if (SW->prn) pprintf("add_Gaussians j=%d lon_Gauss=%f lat_Gauss = %f tt=%e \n", j, RAD2DEG(lon_g), RAD2DEG(lat_g), tt);
        double Q = g[7];
        int isdegsize = (int)g[8];
        double sigma = g[9], radius = g[10], inject = g[11];
        if (isdegsize < 0 || isdegsize > 1) {
            sw2_set_error("Perturbation table error: isdegsize should be 0 or 1");
            return -1;
        }
        if (isdegsize == 1) {
            sigma = sigma / 360.0 * 2 * PI_ * (SW->rE + SW->rP) / 2.0;
            radius = radius / 360.0 * 2 * PI_ * (SW->rE + SW->rP) / 2.0;
        }
        double Qexp = 0, tauTH = 6 * tau;
        double t_int = tau * log(exp((t1g - tauTH) / tau) + exp(t0g / tau));
        if ((t < tauTH + t0g) && (t < t_int)) Qexp = Q * (1.0 - exp(-(t - t0g) / tau));
        if ((t0g + tauTH <= t) && (t <= t1g - tauTH)) Qexp = Q;
        if ((t > t1g - tauTH) && (t > t_int)) Qexp = Q * exp(-(t - (t1g - tauTH)) / tau);
        /* add_Gaussian_core (perturbs.c:205-345) */
        double volume = Dt * Qexp, lon = lon_g, lat = lat_g;
        if (SW->prn) pprintf("add_Gaussian_core lon=%e DEG lat=%e DEG volume=%e sigma=%e radius=%e \n", RAD2DEG(lon), RAD2DEG(lat), volume, sigma, radius);
        while (lon < SW->lon0) lon += (SW->lon1 - SW->lon0);
        while (lon > SW->lon1) lon -= (SW->lon1 - SW->lon0);
        double amp = volume / (2.0 * PI_ * (sigma * sigma) * (1.0 - exp(-(radius * radius) / (2.0 * (sigma * sigma)))));
        if (SW->prn) pprintf("add_Gaussian core lon=%.3f lat=%.3f volume=%.3e amplitude=%.3f\n", RAD2DEG(lon), RAD2DEG(lat), volume, amp);
        double *pf = inc[K];
        memset(pf, 0, sizeof(double) * (size_t)M->S);
        memset(trval[K], 0, sizeof(double) * (size_t)M->S);
        memset(mask[K], 0, (size_t)M->S);
        gaussian_field(SW, lon, lat, amp, sigma, radius, pf, inject, trval[K], mask[K]);
        double lonRight = lon + (SW->lon1 - SW->lon0), lonLeft = lon - (SW->lon1 - SW->lon0);
        if ((lon + radius / g_rZ(SW, lat)) >= SW->lon1) gaussian_field(SW, lonLeft, lat, amp, sigma, radius, pf, inject, trval[K], mask[K]);
        if ((lon - radius / g_rZ(SW, lat)) <= SW->lon0) gaussian_field(SW, lonRight, lat, amp, sigma, radius, pf, inject, trval[K], mask[K]);
        double V = 0.0;
        for (int jj = 1; jj <= SW->ny; jj++)
            for (int ii = 1; ii <= SW->nx; ii++) {
                double a = ac(pf, ii, jj, M);
                double s = ac(SW->Dxc, ii, jj, M) * ac(SW->Dyc, ii, jj, M);
                double v = a * s;
                V += v;
            }
        if (volume == 0.0) pprintf("Volume to be injected is currently 0.\n");
        if (fabs(volume) > 1e-7) {
            if (V == 0.0) {
                sw2_set_error("Global total injected volume Vg=%e equals 0", V);
                return -1;
            }
            double rf = (volume / V), Vcheck = 0.0;
            if (SW->prn) pprintf("Perturbation reescale factor is %e\n", rf);
            for (int jj = 1; jj <= SW->ny; jj++)
                for (int ii = 1; ii <= SW->nx; ii++) {
                    double a = rf * ac(pf, ii, jj, M);
                    ac(pf, ii, jj, M) = a; /* the increment eta receives */
                    double s = ac(SW->Dxc, ii, jj, M) * ac(SW->Dyc, ii, jj, M);
                    double v = a * s;
                    Vcheck += v;
                }
            if (fabs(Vcheck - volume) / volume > 1e-1) {
                sw2_set_error("Total injected volume Vcheck=%e does not match input volume=%e (difference is=%e)", Vcheck, volume, fabs(Vcheck - volume));
                return -1;
            }
        } else {
            /* no eta change: mark the increment as absent */
            pf[0] = NAN;
        }
        K++;
    }
    return K;
}

/* ---- sparse injections: the cells inside the truncation radius only ---- */
typedef struct {
    int q, call;
    double pf, tv;
} inj_cell;

static int inj_cmp(const void *a, const void *b) {
    const inj_cell *x = (const inj_cell *)a, *y = (const inj_cell *)b;
    if (x->q != y->q) return x->q < y->q ? -1 : 1;
    return x->call < y->call ? -1 : (x->call > y->call);
}

typedef struct {
    int i;
    double c;
} col_cos;

/* cells of one row (one call: distinct columns) by field offset.  Short
 * rows: insertion sort.  Longer ones: each cell goes to the slot of its
 * offset within the row's span (slot: scratch of nslot ints, all -1 on entry
 * and on return), read back in order; a span too wide for that is sorted. */
static void sort_row(inj_cell *c, int n, int *slot, int nslot, inj_cell *tmp) {
    if (n <= 16) {
        for (int a = 1; a < n; a++) {
            inj_cell x = c[a];
            int b = a - 1;
            for (; b >= 0 && c[b].q > x.q; b--) c[b + 1] = c[b];
            c[b + 1] = x;
        }
        return;
    }
    int lo = c[0].q, hi = c[0].q;
    for (int a = 1; a < n; a++) {
        if (c[a].q < lo) lo = c[a].q;
        if (c[a].q > hi) hi = c[a].q;
    }
    const int span = hi - lo + 1;
    if (span > nslot || span > 8 * n) {
        qsort(c, (size_t)n, sizeof(inj_cell), inj_cmp);
        return;
    }
    for (int a = 0; a < n; a++) slot[c[a].q - lo] = a;
    int m = 0;
    for (int k = 0; k < span; k++)
        if (slot[k] >= 0) {
            tmp[m++] = c[slot[k]];
            slot[k] = -1;
        }
    memcpy(c, tmp, sizeof(inj_cell) * (size_t)n);
}

static int col_cmp(const void *a, const void *b) {
    const col_cos *x = (const col_cos *)a, *y = (const col_cos *)b;
    if (x->c != y->c) return x->c > y->c ? -1 : 1;
    return x->i - y->i;
}

/* the accepted cells of rows jlo..jhi for one centre (see gaussian_cells),
 * appended to *cells in row order, each row sorted by column */
static int band_cells(sw *SW, int jlo, int jhi, double lat0, double Rm, double amp, double sigma, double radius,
                      double inject, int call, inj_cell **cells, int *n, int *cap, const col_cos *cols, int ntrust) {
    map *M = SW->M;
    const int nslot = SW->nx + 4;
    int *slot = (int *)malloc(sizeof(int) * (size_t)nslot);
    inj_cell *tmp = (inj_cell *)malloc(sizeof(inj_cell) * (size_t)nslot);
    if (!slot || !tmp) {
        free(slot);
        free(tmp);
        return -1;
    }
    for (int k = 0; k < nslot; k++) slot[k] = -1;
    int rc = 0;
    /* the list in locals: threads keep their counters in registers, not in
     * a cache line they share */
    inj_cell *lc = *cells;
    int ln = *n, lcap = *cap;
    for (int j = jlo; j <= jhi && rc == 0; j++) {
        const double lat1 = g_lat2pclat(SW, g_latc(SW, j));
        const double a = sin(lat0) * sin(lat1), b = cos(lat0) * cos(lat1);
        const double rbest = Rm * acos(a + b * cols[0].c);
        if (!(rbest < radius * (1.0 + 1e-9))) continue;
        const int row0 = ln;
        int misses = 0;
        for (int k = 0; misses < 8; k++) {
            if (k >= ntrust) {
                if (ntrust < SW->nx) rc = 1; /* the walk needs columns past the sorted window */
                break;
            }
            const int i = cols[k].i;
            /* great_circle(lon, lat0, lonc(i), lat1, Rm) term by term: its
             * two latitude products are the row's a and b, and its
             * cos(lonc(i) - lon) is cols[k].c, the same call on the same
             * operands */
            double r = Rm * acos(a + b * cols[k].c);
            if (!(r < radius)) {
                misses++;
                continue;
            }
            misses = 0;
            if (ln == lcap) {
                int nc = lcap ? 2 * lcap : 4096;
                inj_cell *nb = (inj_cell *)realloc(lc, sizeof(inj_cell) * (size_t)nc);
                if (!nb) {
                    rc = -1;
                    break;
                }
                lc = nb;
                lcap = nc;
            }
            inj_cell *e = &lc[ln++];
            e->q = LP(i, j, M);
            e->call = call;
            /* SW2 evaluates this exp twice, for eta and for the tracer:
             * one call gives both values */
            const double ex = exp(-(r * r) / (2.0 * (sigma * sigma)));
            e->pf = amp * ex;
            e->tv = (inject) * ex;
        }
        if (rc == 0) sort_row(lc + row0, ln - row0, slot, nslot, tmp);
    }
    *cells = lc;
    *n = ln;
    *cap = lcap;
    free(slot);
    free(tmp);
    return rc;
}

/* the columns in descending cos(lonc - lon) order, as far as a walk may go.
 * With 2H+1 < nx only the 2H+1 columns nearest to lon (cyclically) are
 * evaluated and sorted; *ntrust = 2H+1-16 of them are the global top:
 * every column outside the window is at least 8 column widths farther from
 * lon than those, so its cosine is smaller by far more than any rounding.
 * Otherwise all columns, all trusted. */
static int sorted_cols(sw *SW, double lon, int H, col_cos *cols, int *ntrust) {
    const int nx = SW->nx;
    int m = 0;
    if (2 * H + 1 >= nx) {
        for (int i = 1; i <= nx; i++) {
            cols[m].i = i;
            cols[m++].c = cos(g_lonc(SW, i) - lon);
        }
        *ntrust = nx;
    } else {
        const double dl = (SW->lon1 - SW->lon0) / nx;
        const int ic = (int)floor((lon - SW->lon0) / dl) + 1; /* the column (or its periodic copy) holding lon */
        for (int d = -H; d <= H; d++) {
            const int i = (((ic + d - 1) % nx) + nx) % nx + 1;
            cols[m].i = i;
            cols[m++].c = cos(g_lonc(SW, i) - lon);
        }
        *ntrust = m - 16;
    }
    qsort(cols, (size_t)m, sizeof(col_cos), col_cmp);
    return m;
}

/* the cells get_Gaussian_field (perturbs.c:348-395) accepts for one centre,
 * appended to *cells: SW2's expressions, evaluated only where they can pass.
 * In a row the distance falls as cos(lonc - lon) grows (cos(lat0)cos(lat1)
 * >= 0 and every rounding is monotone), so the accepted columns are a
 * prefix of the columns ordered by that cosine; the walk goes 8 columns
 * past the last acceptance, and rows are candidates when their best column
 * comes within a relative 1e-9 of the radius. */
static int gaussian_cells(sw *SW, double lon, double lat, double amp, double sigma, double radius, double inject, int call,
                          inj_cell **cells, int *n, int *cap, col_cos *cols) {
    const int ny = SW->ny;
    const double Rm = (SW->rE + SW->rP) / 2.0, lat0 = g_lat2pclat(SW, lat);
    /* Rows whose (parametric) latitude lies farther than the radius from
     * the centre's hold no accepted cell: the central angle between two
     * points is at least their latitude difference, and the computed
     * distance is within a relative 1e-6 of the true one once the angle
     * exceeds 1e-4 rad (the error of acos near 1 grows as 1/angle).  The
     * latitude rises with j, so the band is found by bisection; tinier
     * radii scan every row. */
    int jlo = 1, jhi = ny;
    const double band = radius / Rm;
    if (band >= 1e-4) {
        const double lim = band * (1.0 + 1e-6);
        int lo = 1, hi = ny + 1; /* first row with lat1 >= lat0 - lim */
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (g_lat2pclat(SW, g_latc(SW, mid)) >= lat0 - lim) hi = mid;
            else lo = mid + 1;
        }
        jlo = lo;
        lo = 1, hi = ny + 1; /* first row with lat1 > lat0 + lim */
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (g_lat2pclat(SW, g_latc(SW, mid)) > lat0 + lim) hi = mid;
            else lo = mid + 1;
        }
        jhi = lo - 1;
    }
    const int nrows = jhi - jlo + 1;
    int nt = 1;
#ifdef _OPENMP
    /* rows are independent: a large band is split into contiguous chunks,
     * one list per thread, joined in row order (the same cells in the same
     * order as one thread) */
    if (nrows >= 32) {
        nt = omp_get_max_threads();
        if (nt > 8) nt = 8;
        if (nt > nrows / 16) nt = nrows / 16;
        if (nt < 1) nt = 1;
    }
#endif
    /* a window of columns first; wider (finally all of them) if a row's
     * walk reaches past its trusted part */
    const int n0 = *n;
    for (int H = 128;; H *= 2) {
        int ntrust;
        sorted_cols(SW, lon, H, cols, &ntrust);
        int r;
        if (nt == 1) {
            r = band_cells(SW, jlo, jhi, lat0, Rm, amp, sigma, radius, inject, call, cells, n, cap, cols, ntrust);
        } else {
            r = 0;
            inj_cell *tc[8] = {NULL};
            int tn[8] = {0}, tcap[8] = {0}, terr[8] = {0};
            /* chunk id covers rows r0..r1 whatever team the runtime grants
             * (a smaller one under OMP_DYNAMIC, a thread limit or a caller's
             * enclosing parallel region): the loop hands every chunk to some
             * thread, and the lists join in chunk order below */
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 1) num_threads(nt)
#endif
            for (int id = 0; id < nt; id++) {
                const int r0 = jlo + (int)((long)nrows * id / nt), r1 = jlo + (int)((long)nrows * (id + 1) / nt) - 1;
                terr[id] = band_cells(SW, r0, r1, lat0, Rm, amp, sigma, radius, inject, call, &tc[id], &tn[id], &tcap[id],
                                      cols, ntrust);
            }
            int total = *n;
            for (int t = 0; t < nt; t++) {
                if (terr[t] < 0) r = -1;
                else if (terr[t] > 0 && r == 0) r = 1;
                total += tn[t];
            }
            if (r == 0 && total > *cap) {
                inj_cell *nb = (inj_cell *)realloc(*cells, sizeof(inj_cell) * (size_t)total);
                if (nb) {
                    *cells = nb;
                    *cap = total;
                } else
                    r = -1;
            }
            for (int t = 0; t < nt; t++) {
                if (r == 0 && tn[t]) {
                    memcpy(*cells + *n, tc[t], sizeof(inj_cell) * (size_t)tn[t]);
                    *n += tn[t];
                }
                free(tc[t]);
            }
        }
        if (r <= 0) return r;
        *n = n0; /* the window was too narrow: again, wider */
    }
}

/* merge the cell lists of the calls (each in field order, calls in order)
 * into field order, a cell of several calls in call order */
static int merge_calls(const inj_cell *c, const int *end, int ncall, inj_cell *out) {
    int at[3] = {0, end[0], end[1]}, m = 0;
    for (;;) {
        int best = -1;
        for (int k = 0; k < ncall; k++) {
            if (at[k] >= end[k]) continue;
            if (best < 0 || c[at[k]].q < c[at[best]].q) best = k;
        }
        if (best < 0) return m;
        out[m++] = c[at[best]++];
    }
}

static int inj_reserve(sw2_inj *x, int n) {
    if (n <= x->cap) return 0;
    int nc = n + n / 2 + 1024;
    int *ni = (int *)realloc(x->idx, sizeof(int) * (size_t)nc);
    if (ni) x->idx = ni;
    double *na = (double *)realloc(x->inc, sizeof(double) * (size_t)nc);
    if (na) x->inc = na;
    double *nt = (double *)realloc(x->trv, sizeof(double) * (size_t)nc);
    if (nt) x->trv = nt;
    if (!ni || !na || !nt) return -1;
    x->cap = nc;
    return 0;
}

void sw2_inj_free(sw2_inj *inj, int k) {
    for (int q = 0; q < k; q++) {
        free(inj[q].idx);
        free(inj[q].inc);
        free(inj[q].trv);
        inj[q] = (sw2_inj){0, 0, 0, NULL, NULL, NULL};
    }
}

int sw2_gaussians_sparse(sw *SW, double Dt, double t, int maxk, sw2_inj *inj) {
    int K = 0;
    col_cos *cols = (col_cos *)malloc(sizeof(col_cos) * (size_t)SW->nx);
    inj_cell *cells = NULL, *merged = NULL;
    int cap = 0, mcap = 0, rc = 0;
    if (!cols) {
        sw2_set_error("out of host memory");
        return -1;
    }
    for (int j = 0; j <= SW->num_Gaussians - 1 && !rc; j++) {
        double *g = SW->Gaussians + j * N_COLS_GAUSS;
        double t0g = g[0], t1g = g[1];
        if (!((t0g <= t) && (t1g >= t))) continue;
        if (K >= maxk) {
            sw2_set_error("more than %d simultaneous injections", maxk);
            rc = -1;
            break;
        }
        /* the scalars of add_Gaussians / add_Gaussian_core, as in
         * sw2_gaussians_at */
        double tau = g[2];
        if (tau < 0) {
            sw2_set_error("Uhhh tau is %e and should be positive", tau);
            rc = -1;
            break;
        }
        double lon_g = DEG2RAD(g[3]), lat_g = DEG2RAD(g[4]);
        int isdegvel = (int)g[5];
        double vel = g[6];
        if (isdegvel < 0 || isdegvel > 1) {
            sw2_set_error("Perturbation table error: isdegvel should be 0 or 1");
            rc = -1;
            break;
        }
        double tt = (t - t0g), Om = 0;
        if (isdegvel == 0) Om = vel / g_rZ(SW, lat_g);
        if (isdegvel == 1) Om = DEG2RAD(vel);
        double lon_g0 = lon_g;
        lon_g = lon_g0 + Om * tt;
        if (SW->prn) pprintf("add_Gaussians j=%d lon_Gauss=%f lat_Gauss = %f tt=%e \n", j, RAD2DEG(lon_g), RAD2DEG(lat_g), tt);
        double Q = g[7];
        int isdegsize = (int)g[8];
        double sigma = g[9], radius = g[10], inject = g[11];
        if (isdegsize < 0 || isdegsize > 1) {
            sw2_set_error("Perturbation table error: isdegsize should be 0 or 1");
            rc = -1;
            break;
        }
        if (isdegsize == 1) {
            sigma = sigma / 360.0 * 2 * PI_ * (SW->rE + SW->rP) / 2.0;
            radius = radius / 360.0 * 2 * PI_ * (SW->rE + SW->rP) / 2.0;
        }
        double Qexp = 0, tauTH = 6 * tau;
        double t_int = tau * log(exp((t1g - tauTH) / tau) + exp(t0g / tau));
        if ((t < tauTH + t0g) && (t < t_int)) Qexp = Q * (1.0 - exp(-(t - t0g) / tau));
        if ((t0g + tauTH <= t) && (t <= t1g - tauTH)) Qexp = Q;
        if ((t > t1g - tauTH) && (t > t_int)) Qexp = Q * exp(-(t - (t1g - tauTH)) / tau);
        double volume = Dt * Qexp, lon = lon_g, lat = lat_g;
        if (SW->prn) pprintf("add_Gaussian_core lon=%e DEG lat=%e DEG volume=%e sigma=%e radius=%e \n", RAD2DEG(lon), RAD2DEG(lat), volume, sigma, radius);
        while (lon < SW->lon0) lon += (SW->lon1 - SW->lon0);
        while (lon > SW->lon1) lon -= (SW->lon1 - SW->lon0);
        double amp = volume / (2.0 * PI_ * (sigma * sigma) * (1.0 - exp(-(radius * radius) / (2.0 * (sigma * sigma)))));
        if (SW->prn) pprintf("add_Gaussian core lon=%.3f lat=%.3f volume=%.3e amplitude=%.3f\n", RAD2DEG(lon), RAD2DEG(lat), volume, amp);
        int n = 0, end[3], ncall = 0;
        if (SW->prn) pprintf("get_Gaussian_field entering lon=%f DEG lat=%f DEG\n", RAD2DEG(lon), RAD2DEG(lat));
        rc |= gaussian_cells(SW, lon, lat, amp, sigma, radius, inject, 0, &cells, &n, &cap, cols);
        end[ncall++] = n;
        double lonRight = lon + (SW->lon1 - SW->lon0), lonLeft = lon - (SW->lon1 - SW->lon0);
        if ((lon + radius / g_rZ(SW, lat)) >= SW->lon1) {
            if (SW->prn) pprintf("get_Gaussian_field entering lon=%f DEG lat=%f DEG\n", RAD2DEG(lonLeft), RAD2DEG(lat));
            rc |= gaussian_cells(SW, lonLeft, lat, amp, sigma, radius, inject, 1, &cells, &n, &cap, cols);
            end[ncall++] = n;
        }
        if ((lon - radius / g_rZ(SW, lat)) <= SW->lon0) {
            if (SW->prn) pprintf("get_Gaussian_field entering lon=%f DEG lat=%f DEG\n", RAD2DEG(lonRight), RAD2DEG(lat));
            rc |= gaussian_cells(SW, lonRight, lat, amp, sigma, radius, inject, 2, &cells, &n, &cap, cols);
            end[ncall++] = n;
        }
        if (!rc && n > mcap) {
            inj_cell *nb = (inj_cell *)realloc(merged, sizeof(inj_cell) * (size_t)n);
            if (nb) {
                merged = nb;
                mcap = n;
            } else
                rc = -1;
        }
        if (rc) {
            sw2_set_error("out of host memory");
            break;
        }
        /* one entry per cell in SW2's row-major order; a cell two centres
         * reach adds their values in call order, the tracer keeps the last.
         * Each call's cells are in field order already (rows in order, each
         * row sorted), so a merge orders them. */
        if (merge_calls(cells, end, ncall, merged) != n) {
            sw2_set_error("injection cell merge lost cells");
            rc = -1;
            break;
        }
        {
            inj_cell *tmp = cells;
            int tcap = cap;
            cells = merged;
            cap = mcap;
            merged = tmp;
            mcap = tcap;
        }
        sw2_inj *x = &inj[K];
        if (inj_reserve(x, n)) {
            sw2_set_error("out of host memory");
            rc = -1;
            break;
        }
        int m = 0;
        for (int a = 0; a < n;) {
            double pf = 0.0, tv = 0.0;
            int b = a;
            for (; b < n && cells[b].q == cells[a].q; b++) {
                pf += cells[b].pf;
                tv = cells[b].tv;
            }
            x->idx[m] = cells[a].q;
            x->inc[m] = pf;
            x->trv[m] = tv;
            m++;
            a = b;
        }
        x->n = m;
        /* the two volume sums over every interior cell in row-major order:
         * cells outside the radius add an exact zero */
        double V = 0.0;
        for (int a = 0; a < m; a++) {
            double s = SW->Dxc[x->idx[a]] * SW->Dyc[x->idx[a]];
            double v = x->inc[a] * s;
            V += v;
        }
        if (volume == 0.0) pprintf("Volume to be injected is currently 0.\n");
        if (fabs(volume) > 1e-7) {
            if (V == 0.0) {
                sw2_set_error("Global total injected volume Vg=%e equals 0", V);
                rc = -1;
                break;
            }
            double rf = (volume / V), Vcheck = 0.0;
            if (SW->prn) pprintf("Perturbation reescale factor is %e\n", rf);
            for (int a = 0; a < m; a++) {
                double v0 = rf * x->inc[a];
                x->inc[a] = v0;
                double s = SW->Dxc[x->idx[a]] * SW->Dyc[x->idx[a]];
                double v = v0 * s;
                Vcheck += v;
            }
            if (fabs(Vcheck - volume) / volume > 1e-1) {
                sw2_set_error("Total injected volume Vcheck=%e does not match input volume=%e (difference is=%e)", Vcheck, volume, fabs(Vcheck - volume));
                rc = -1;
                break;
            }
            x->has_inc = 1;
        } else {
            x->has_inc = 0;
        }
        K++;
    }
    free(cols);
    free(cells);
    free(merged);
    return rc ? -1 : K;
}

int sw2_vortices_at(sw *SW, double Dt, double t, int maxk, double **deta, double **du, double **dv) {
    int K = 0;
    map *M = SW->M;
    for (int k = 0; k <= SW->num_vortices - 1; k++) {
        double *w = SW->vortices + k * N_COLS_VORTICES;
        double t0v = w[0];
        if (!((t0v < t) && (t0v >= (t - Dt)))) continue;
        if (K >= maxk) {
            sw2_set_error("more than %d simultaneous vortices", maxk);
            return -1;
        }
        double lonc = w[1], latc = w[2], as = w[3], bs = w[4], Umax = w[5], n = w[6];
        double amp = 0.5 * Umax * g_rM(SW, DEG2RAD(latc)) * DEG2RAD(bs) * g_f(SW, DEG2RAD(latc)) *
                     pow((1.0 - 1.0 / (2.0 * n)), ((1.0 - 2.0 * n) / (2.0 * n))) * exp(1.0 - 1.0 / (2.0 * n)) /
                     (g_geff(SW, DEG2RAD(latc)) * n);
        pprintf("add_vortices t0_vortex=%f, lonc=%f DEG, latc=%f DEG, as=%f DEG, bs=%f DEG, amp=%e, Umax=%f, n=%f\n", t0v, lonc, latc, as, bs, amp, Umax, n);
        memset(deta[K], 0, sizeof(double) * (size_t)M->S);
        memset(du[K], 0, sizeof(double) * (size_t)M->S);
        memset(dv[K], 0, sizeof(double) * (size_t)M->S);
        for (int j = 1; j <= SW->ny; j++)
            for (int i = 1; i <= SW->nx; i++) {
                double xa = (RAD2DEG(g_lonc(SW, i)) - lonc) / as, ya = (RAD2DEG(g_latc(SW, j)) - latc) / bs;
                double ex = pow(xa * xa + ya * ya, n);
                double eij = amp * exp(-ex);
                ac(deta[K], i, j, M) = eij;
                double xu = (RAD2DEG(g_lone(SW, i)) - lonc) / as;
                double pu = xu * xu + ya * ya;
                ac(du[K], i, j, M) = 2.0 * (n)*ac(SW->ge, i, j, M) / (g_f(SW, DEG2RAD(latc)) * (DEG2RAD(bs) * DEG2RAD(bs)) * g_rM(SW, DEG2RAD(latc))) *
                                     (g_latc(SW, j) - DEG2RAD(latc)) * pow(pu, n - 1.0) * eij;
                double yv = (RAD2DEG(g_latn(SW, j)) - latc) / bs;
                double pv = xa * xa + yv * yv;
                ac(dv[K], i, j, M) = -2.0 * (n)*ac(SW->gn, i, j, M) / (g_f(SW, DEG2RAD(latc)) * (DEG2RAD(as) * DEG2RAD(as)) * g_rZ(SW, DEG2RAD(latc))) *
                                     (g_lonc(SW, i) - DEG2RAD(lonc)) * pow(pv, n - 1.0) * eij;
            }
        pprintf("add_vortex_core lonc=%f DEG, latc=%f DEG, as=%f DEG, bs=%f DEG\n", lonc, latc, as, bs);
        K++;
    }
    return K;
}

/* host pre-loop perturbations applied directly to SW (core.c:94-96) */
static void host_perturb(sw *SW, double Dt, double t) {
    map *M = SW->M;
    int S = M->S;
    double *inc[8], *trv[8], *dd[3][8];
    unsigned char *mk[8];
    for (int k = 0; k < 8; k++) {
        inc[k] = (double *)malloc(sizeof(double) * S);
        trv[k] = (double *)malloc(sizeof(double) * S);
        mk[k] = (unsigned char *)malloc(S);
        for (int q = 0; q < 3; q++) dd[q][k] = (double *)malloc(sizeof(double) * S);
    }
    int K = sw2_gaussians_at(SW, Dt, t, 8, inc, trv, mk);
    if (K < 0) CRASH("%s", sw2_last_error());
    for (int k = 0; k < K; k++) {
        for (int q = 0; q < S; q++)
            if (mk[k][q]) {
                SW->tracer[q] = trv[k][q];
                SW->perturbs[q] = 1.0;
            }
        if (!isnan(inc[k][0]))
            for (int j = 1; j <= SW->ny; j++)
                for (int i = 1; i <= SW->nx; i++) ac(SW->eta, i, j, M) += ac(inc[k], i, j, M);
    }
    int V = sw2_vortices_at(SW, Dt, t, 8, dd[0], dd[1], dd[2]);
    if (V < 0) CRASH("%s", sw2_last_error());
    for (int k = 0; k < V; k++) {
        for (int j = 1; j <= SW->ny; j++)
            for (int i = 1; i <= SW->nx; i++) {
                ac(SW->eta, i, j, M) += ac(dd[0][k], i, j, M);
                ac(SW->u, i, j, M) += ac(dd[1][k], i, j, M);
                ac(SW->v, i, j, M) += ac(dd[2][k], i, j, M);
            }
        halo_update(SW->eta, M);
        halo_update(SW->u, M);
        halo_update(SW->v, M);
    }
    for (int k = 0; k < 8; k++) {
        free(inc[k]);
        free(trv[k]);
        free(mk[k]);
        for (int q = 0; q < 3; q++) free(dd[q][k]);
    }
}

/* ------------------------------------------------------------------ loop */
void loop(char *bname, char *output_folder, sw *SW) {
    map *M = SW->M;
    double t = SW->t0, Dt = SW->Dt, ke, ape;
    pprintf("Auxiliary matrices allocated.\n");
    halo_update(SW->tracer, M); /* core.c:88-101 */
    halo_update(SW->u, M);
    halo_update(SW->v, M);
    halo_update(SW->eta, M);
    halo_update(SW->hB, M);
    memset(SW->perturbs, 0, sizeof(double) * (size_t)M->S);
    int prn_saved = SW->prn;
    SW->prn = 0;
    host_perturb(SW, Dt, t);
    SW->prn = prn_saved;
    halo_update(SW->tracer, M);
    halo_update(SW->eta, M);
    do_channel(SW);
    pprintf("Halos updated.\n");
    cr_start("sim_time", 0);
    stats_line(SW, SW->tracer, "tracer");
    stats_line(SW, SW->eta, "eta");
    stats_line(SW, SW->u, "u");
    stats_line(SW, SW->v, "v");
    int frame = 0, nstart;
    if (SW->loadFrom != 0) {
        frame = SW->loadFrom;
        if (load_binary_sw(frame, &nstart, &t, bname, output_folder, SW)) CRASH("File to be loaded not found");
    } else {
        nstart = 0;
        save_binary_sw(frame, 0, t, bname, output_folder, SW);
    }
    calc_energy_balance(&ke, &ape, SW);
    print_energy(t, ke, ape);
    frame++;
    nstart++;

    sw2_session *s = NULL;
    if (sw2_session_create(SW, SW2_FORM_FP64, &s)) CRASH("%s", sw2_last_error());
    if (sw2_session_begin(s, nstart - 1, t)) CRASH("%s", sw2_last_error());
    char msg[512];
    int n = nstart;
    while (n <= SW->N) {
        /* steps that need the host: diagnostics (core.c:136) and saves */
        int e = n;
        while (e <= SW->N && !(e == nstart || e % SW->IteInfo == 0 || e % SW->saveEvery == 0)) e++;
        int quiet = (e > SW->N ? SW->N : e - 1) - n + 1;
        if (quiet > 0) {
            SW->prn = 0;
            cr_start("time_step", 0);
            if (sw2_session_step(s, quiet)) CRASH("%s", sw2_last_error());
            cr_end("time_step", 0);
        }
        if (sw2_session_failed_step(s, msg, sizeof msg)) {
            sw2_session_download(s, SW);
            sw2_session_destroy(s);
            CRASH("%s", msg);
        }
        if (e > SW->N) break;
        int cur;
        double tcur;
        sw2_session_state(s, &cur, &tcur);
        SW->prn = (e == nstart || e % SW->IteInfo == 0);
        cr_start("time_step", 0);
        if (SW->prn) pprintf("===SW STARTING TIME STEP %d / %d from t0=%e to t=%e \n", e, SW->N, tcur, tcur + Dt);
        if (sw2_session_step(s, 1)) CRASH("%s", sw2_last_error());
        if (sw2_session_failed_step(s, msg, sizeof msg)) {
            sw2_session_download(s, SW);
            sw2_session_destroy(s);
            CRASH("%s", msg);
        }
        sw2_session_state(s, &cur, &tcur);
        if (SW->prn || e % SW->saveEvery == 0) sw2_session_download(s, SW);
        if (SW->prn) {
            double dtcfl = 0;
            sw2_session_last_cfl(s, &dtcfl);
            pprintf("Smallest possible timestep is Dt_CFL=%.6e while Dt=%.6e.\n", dtcfl, Dt);
            const char *names[4] = {"tracer", "eta", "u", "v"};
            double *fl[4] = {SW->tracer, SW->eta, SW->u, SW->v};
            for (int q = 0; q < 4; q++) {
                double a, b, c;
                get_stats(SW, fl[q], &a, &b, &c);
                pprintf("stats_scaf fname=<%s> step=%d min=%e max=%e avg=%e \n", names[q], e, a, b, c);
            }
        }
        if (e % SW->saveEvery == 0) {
            save_binary_sw(frame, e, tcur, bname, output_folder, SW);
            frame++;
        }
        if (SW->prn) {
            calc_energy_balance(&ke, &ape, SW);
            print_energy(tcur, ke, ape);
            pprintf("===TIME STEP END TIME STEP %d wall clock seconds per timestep %.3f = %.3f ms; TOTAL per timestep= %.3f ms\n", e,
                    cr_time("time_step", 0), cr_time("time_step", 0) * 1000, cr_time("time_step", 0) * 1000 * quants());
        }
        cr_end("time_step", 0);
        n = e + 1;
    }
    sw2_session_download(s, SW);
    sw2_session_destroy(s);
    cr_end("sim_time", 0);
    pprintf("Core loop ended.\n");
    pprintf("Memory for auxiliary variables freed.\n");
}
