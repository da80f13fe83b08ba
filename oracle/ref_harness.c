/*
 * ref_harness.c -- drives SW2's own code, compiled from the reference's
 * sources into oracle/_ref (oracle/Makefile, target `ref`), so that the
 * oracle can be checked against the real SW2 function by function.
 *
 * TEST INFRASTRUCTURE ONLY: tests/ and bench.py's cpu_baseline leg use it as
 * a checker and as the timed CPU baseline; the product never links or calls
 * it.  Built with gcc against the MPICH of this image (MPI_Init as a
 * singleton in a test process, or under mpiexec for the multi-rank timing).
 *
 * Two departures from SW2's main/Main.c, both outside the arithmetic:
 *   - SW2 reads its case through p_startinput (sppde/sppde_parser.c:177-210),
 *     which first runs m4 over the file; this image has no m4.  The harness
 *     opens a case file that holds no m4 construct and reads it with SW2's
 *     own tag readers (p_getint, p_getdouble, p_getndoubles and the table
 *     readers of io.c), tag by tag in the order of init.c:48-229; every other
 *     step of the set-up (createM, init_channel_region, init_coords_spheroid,
 *     read_vortices / read_Gaussians / read_wind, the halo updates) is SW2's.
 *   - SW2 reads memory it never wrote: windy (init.c:204,234) and the psi /
 *     xi halos of laplacian_eta (laplacian.c:117-127).  In SW2's own runs of
 *     grids above glibc's mmap threshold those reads see fresh zero pages.
 *     The harness makes that hold for every size and every state of the
 *     process's heap: the library defines malloc as calloc (below), and
 *     -Bsymbolic binds SW2's calls to that definition.  (Mapping every block
 *     of 4 KiB or more with mallopt was not enough: a request that a freed
 *     chunk of the heap can serve is served from it, with whatever an
 *     earlier user left there, and in a test process that held SW2's reads
 *     to the history of the process.)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpi.h"
#include "sppde.h"
#include "sppde_parser.h"
#include "sw.h"

static int g_started = 0;

/* SW2's allocations, zero-filled (see above).  Only calls made from inside
 * this library bind here (-Bsymbolic; the library is not loaded into the
 * global scope); the C library's own calls keep its malloc. */
void *malloc(size_t n) { return calloc(1, n); }

/* MPI (a singleton unless started by mpiexec) and SW2's output channel */
int ref_start(const char *outdir) {
    int on = 0;
    if (g_started) return 0;
    MPI_Initialized(&on);
    if (!on) checkr(MPI_Init(NULL, NULL), "init");
    /* mode 2: rank 0 prints to stdout, the others print nothing */
    init_pprintf(2, 0, (char *)outdir);
    cr_reset();
    g_started = 1;
    return 0;
}

void ref_stop(void) {
    if (!g_started) return;
    end_pprintf();
    MPI_Finalize();
    g_started = 0;
}

/* The world of a case file without m4 constructs, on npx x npy ranks
 * (channel geometry only: polar must be 0).  Tags in the order of
 * init.c:48-229. */
sw *ref_world(int npx, int npy, const char *casefile) {
    FILE *fin = fopen(casefile, "r");
    if (!fin) CRASH("ref_world: cannot open <%s>", casefile);
    sw *SW = (sw *)calloc(1, sizeof(sw));
    p_setdebug(0);
    SW->sch = p_getint(fin, "sch");
    SW->polar = p_getint(fin, "polar");
    if (SW->polar != 0) CRASH("ref_world: channel cases only (polar=%d)", SW->polar);
    SW->lon0 = deg2rad(p_getdouble(fin, "lon0"));
    SW->lon1 = deg2rad(p_getdouble(fin, "lon1"));
    SW->lat0 = deg2rad(p_getdouble(fin, "lat0"));
    SW->lat1 = deg2rad(p_getdouble(fin, "lat1"));
    SW->nx = p_getint(fin, "nx");
    SW->ny = p_getint(fin, "ny");
    SW->nz = 1;
    if (npx * npy != quants()) CRASH("ref_world: %d x %d ranks asked, %d running", npx, npy, quants());
    {
        int gl0[4] = {0, 1, 1, 0}, gl1[4] = {0, SW->nx, SW->ny, SW->nz};
        int np[4] = {0, npx, npy, 1}, per[4] = {0, IS_PERIODIC_X, 0, 0};
        SW->M = &(SW->m);
        createM(2, gl0, gl1, np, 2, per, SW->M);
    }
    SW->t0 = p_getdouble(fin, "t0");
    SW->t1 = p_getdouble(fin, "t1");
    SW->Dt = p_getdouble(fin, "Dt");
    SW->Courant_max = p_getdouble(fin, "CFL");
    SW->N = (int)(((SW->t1) - (SW->t0)) / (SW->Dt));
    {
        double tau = p_getdouble(fin, "Tau");
        SW->sigma_dis = tau > 1e10 ? 0 : 1.0 / tau;
    }
    {
        double nu[3];
        p_getndoubles(fin, 3, nu, "Hypernu");
        SW->nu2 = nu[0];
        SW->nu4 = -nu[1];
        SW->nu6 = nu[2];
    }
    SW->IteInfo = p_getint(fin, "IteInfo");
    SW->saveEvery = p_getint(fin, "SaveEvery");
    SW->loadFrom = p_getint(fin, "LoadFrom");
    SW->reftype = p_getint(fin, "reftype");
    if (SW->reftype != 0) CRASH("ref_world: reftype 0 only");
    SW->reftable = NULL;
    SW->refname[0] = '\0';
    SW->Omega = p_getdouble(fin, "Omega");
    SW->rE = p_getdouble(fin, "rE");
    SW->rP = p_getdouble(fin, "rP");
    SW->epsilon2 = POW2((SW->rE) / (SW->rP));
    SW->gE = p_getdouble(fin, "gE");
    SW->depth = p_getdouble(fin, "depth");
    SW->beta_sponge = 1e5;
    SW->sigma_sponge = 0;
    SW->tau_sponge = 0;
    double **geo[] = {&SW->Dxe, &SW->Dxc, &SW->Dxv, &SW->Dyn, &SW->Dyc, &SW->Dyv, &SW->xe, &SW->xc, &SW->xv, &SW->yn,
                      &SW->yc, &SW->yv, &SW->ge, &SW->gn, &SW->gc, &SW->gv, &SW->fe, &SW->fn, &SW->fc, &SW->fv};
    for (size_t k = 0; k < sizeof(geo) / sizeof(geo[0]); k++) *geo[k] = dmem(SW->M);
    init_channel_region(SW);
    init_coords_spheroid(SW);
    double **fld[] = {&SW->tracer, &SW->perturbs, &SW->u, &SW->v, &SW->windx, &SW->windy, &SW->eta, &SW->hB};
    for (size_t k = 0; k < sizeof(fld) / sizeof(fld[0]); k++) *fld[k] = dmem(SW->M);
    setzero_scaf(SW->tracer, SW->M);
    setzero_scaf(SW->perturbs, SW->M);
    setzero_scaf(SW->u, SW->M);
    setzero_scaf(SW->v, SW->M);
    setzero_scaf(SW->eta, SW->M);
    setzero_scaf(SW->hB, SW->M);
    SW->geoeq = p_getint(fin, "geoeq");
    read_vortices(fin, SW);
    read_Gaussians(fin, SW);
    SW->tracer_dis = p_getdouble(fin, "TracerDis");
    read_wind(fin, SW);
    halo_update(SW->windx, SW->M);
    halo_update(SW->windy, SW->M);
    fclose(fin);
    return SW;
}

void ref_loop(sw *SW, const char *bname, const char *outdir) { loop((char *)bname, (char *)outdir, SW); }

void ref_free(sw *SW) {
    destroy(SW);
    free(SW);
}

/* field access for the binding: 0 eta, 1 u, 2 v, 3 tracer, 4 perturbs,
 * 5 windx, 6 windy, 7 hB; row tables 10.. Dxe Dxc Dxv Dyn Dyc Dyv ge gn gc
 * gv fe fn fc fv (2-D arrays whose values depend on j only) */
double *ref_field(sw *SW, int which) {
    double *f[] = {SW->eta, SW->u, SW->v, SW->tracer, SW->perturbs, SW->windx, SW->windy, SW->hB};
    double *r[] = {SW->Dxe, SW->Dxc, SW->Dxv, SW->Dyn, SW->Dyc, SW->Dyv, SW->ge,
                   SW->gn,  SW->gc,  SW->gv,  SW->fe,  SW->fn,  SW->fc,  SW->fv};
    if (which >= 0 && which < 8) return f[which];
    if (which >= 10 && which < 24) return r[which - 10];
    return NULL;
}

int ref_size(sw *SW) { return SW->M->S; }
int ref_nsteps(sw *SW) { return SW->N; }
/* owned block of this rank: l0/l1 per axis (1 x, 2 y) */
void ref_block(sw *SW, int *out4) {
    out4[0] = SW->M->l0[1];
    out4[1] = SW->M->l1[1];
    out4[2] = SW->M->l0[2];
    out4[3] = SW->M->l1[2];
}

/* the reference's table interpolation (tools.c:21-74) on a table of n points */
double ref_linterp(double x, double *xs, double *ys, int n) { return linterp(x, xs, ys, n); }
