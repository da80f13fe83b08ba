/*
 * sw2_hip.h -- C-ABI of the MI355X (gfx950) shallow-water step library.
 *
 * The library is a drop-in for the time loop of SW2 (Shallow Worlds 2): the
 * same three entry points that SW2's driver calls (main/Main.c:72-78), the
 * same field layout (sppde map with a halo of 2, x fastest), and the same
 * case-file input.  Everything that SW2 computes per step inside loop()
 * (core.c:134-338) runs on the GPU; set-up, case parsing, frame IO and the
 * diagnostics printed on the host stay in host C.
 *
 * Three layers are exported:
 *   1. the coarse SW2 boundary: initialize / loop / destroy (sw.h:222-224);
 *   2. the sppde helpers that SW2's driver links against (sppde/sppde.h);
 *   3. the sw2_* extension API: building a world from parameters instead of a
 *      case file, device-resident sessions (the step loop without host IO),
 *      diagnostics, and the batched / served ensemble API.
 *
 * All signatures use plain C types and pointers.  Fields are host arrays of
 * M->S doubles laid out as sppde does (ac(v,i,j,M) below).
 */
#ifndef SW2_HIP_H
#define SW2_HIP_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * sppde map and field access (sppde/sppde.h:44-67, 131-142), field for field
 * as SW2 declares it, so that code compiled against SW2's own headers reads
 * every member of map and sw at the offsets this library writes.  Only one
 * map is ever in use: one GPU owns the whole domain, x is periodic, y is
 * walled (channel).  The per-process tables al0 / al1 / pco hold entry 0.
 * ------------------------------------------------------------------------- */
#define SW2_MAXP 1920 /* sppde.h:44 (MAXP): slots of the per-process tables */

typedef struct {
    int sh;        /* halo width (2) */
    int nd;        /* number of dimensions (2) */
    int l0[4];     /* first owned index per axis (1-based axes) */
    int l1[4];     /* last owned index per axis */
    int gl0[4];    /* global first index */
    int gl1[4];    /* global last index */
    int al0[SW2_MAXP][4]; /* first index per process and axis (process 0 only) */
    int al1[SW2_MAXP][4]; /* last index per process and axis */
    int pco[SW2_MAXP][4]; /* processor coordinates per process */
    int s[4];      /* extent per axis including halos */
    int S;         /* number of doubles in a field */
    int gpc[4];    /* processors per axis (all 1) */
    int pc[4];     /* my processor coordinates (all 1) */
    int nb_e, nb_w, nb_n, nb_s; /* neighbours: 0 (myself) along periodic x, -1 along y */
    int nb_t, nb_b;             /* top / bottom (3-D maps only): -1 */
    int bc0[4];    /* do I own the initial boundary of axis a (0 if periodic) */
    int bc1[4];    /* do I own the final boundary of axis a (0 if periodic) */
    int per[4];    /* periodicity per axis */
} map;

/* position of (i,j) in a field: x fastest, halo of sh cells on each side */
#define LP(i, j, M) ((i) - ((M)->l0[1]) + (M)->sh + ((j) - ((M)->l0[2]) + (M)->sh) * ((M)->s[1]))
#define ac(v, i, j, M) *((v) + LP(i, j, M))

#ifndef MAXOUTLEN
#define MAXOUTLEN 1000
#endif

/* version string the driver prints (sw.h:13 holds SW2's) */
#ifndef VERSION
#define VERSION "SW 0.125 (gfx950)"
#endif

/* ---------------------------------------------------------------------------
 * The world (sw.h:22-95).  Field names, types and meaning are SW2's so that
 * a driver written against sw.h compiles unchanged against this header.
 * ------------------------------------------------------------------------- */
typedef struct {
    int sch;                 /* limiter: 0 MUSCL .. 9 Lax-Wendroff (calc.c:54-63) */
    int polar;               /* 0: channel (the only geometry on the GPU path) */
    int N;                   /* number of steps */
    int loadFrom;            /* frame to start from (0: none) */
    int saveEvery;           /* frame period in steps */
    int IteInfo;             /* diagnostics period in steps */
    int prn;                 /* diagnostics on this step */
    double t0, t1, Dt;
    double Courant_max;
    int nx, ny, nz;
    double lon0, lat0;       /* radians */
    double lon1, lat1;       /* radians */
    map m;
    map *M;

    double Omega;
    double rE, rP;
    double epsilon2;
    double gE;
    double depth;
    double sigma_dis;
    double tracer_dis;

    double beta_sponge;
    double sigma_sponge;
    double tau_sponge;

    double *windx;
    double *windy;
    double *u, *v, *eta, *hB;
    double *tracer, *perturbs;
    double *xe, *xc, *xv;
    double *yn, *yc, *yv;
    double *Dxe, *Dxc, *Dxv, *Dyn, *Dyc, *Dyv;
    double *ge, *gn, *gc, *gv, *fe, *fn, *fc, *fv;

    double nu2, nu4, nu6;

    int reftype;
    char refname[MAXOUTLEN];
    double *reftable;
    int reftable_nr, reftable_nc;

    int geoeq;

    int num_Gaussians;
    double *Gaussians;       /* num_Gaussians x 12, row major (perturbs.c:55-56) */

    int num_vortices;
    double *vortices;        /* num_vortices x 7 */
} sw;

#define N_COLS_GAUSS 12
#define N_COLS_VORTICES 7

/* ---------------------------------------------------------------------------
 * 1. Coarse SW2 boundary.
 * ------------------------------------------------------------------------- */

/* Replaces initialize() of init.c:23 (sw.h:222).  Reads <bname>.sw (m4
 * define/include subset, <expr> evaluation, tag order of init.c:48-229) and
 * builds the world.  npx*npy must be 1 unless force1proc==1 (one GPU owns the
 * whole domain).  Returns 0; fatal errors abort like SW2's CRASH. */
int initialize(int npx, int npy, int force1proc, char *bname, int copyinput, sw *SW);

/* Replaces loop() of core.c:22 (sw.h:223).  Runs steps nstart+1..N on the
 * GPU with SW2's per-step order and halo semantics, writes the frames and the
 * ENERGY / stats lines SW2 writes, and leaves eta, u, v, tracer and perturbs
 * (halos included) holding the final state in host memory. */
void loop(char *bname, char *output_folder, sw *SW);

/* Replaces destroy() of finish.c:21 (sw.h:224). */
void destroy(sw *SW);

/* ---------------------------------------------------------------------------
 * 2. sppde helpers the SW2 driver calls (sppde/sppde.h:21-40, 353-362).
 *    Single process: quants() is 1 and pprintf writes to stdout.
 * ------------------------------------------------------------------------- */
void checkr(int r, char *s);
void crash(int line, char *file, char *fmt, ...);
#define CRASH(...) crash(__LINE__, __FILE__, __VA_ARGS__)
void init_pprintf(int mode, int flushall, char *output_folder);
void end_pprintf(void);
int pprintf(char *fmt, ...);
int quants(void);
int quisoc(void);
double *dmem(map *M);
void cr_reset(void);
void cr_start(char *s, int level);
void cr_end(char *s, int level);
double cr_time(char *s, int level);
void cr_info(void);

/* ---------------------------------------------------------------------------
 * 3. sw2_* extension API (no counterpart in SW2 unless stated).
 * ------------------------------------------------------------------------- */

/* Case parameters: what initialize() reads from a case file (init.c:48-229),
 * angles in degrees as written in the file.  Tables are borrowed pointers. */
typedef struct {
    int sch, polar;
    double lon0_deg, lon1_deg, lat0_deg, lat1_deg;
    int nx, ny;
    double t0, t1, Dt, CFL;
    double Tau;              /* >1e10 disables the Rayleigh drag */
    double hypernu[3];       /* as written: nu2, nu4 (entered positive), nu6 */
    int IteInfo, SaveEvery, LoadFrom, reftype;
    double Omega, rE, rP, gE, depth;
    int geoeq;
    int num_vortices;  const double *vortices;   /* rows x 7 */
    int num_Gaussians; const double *Gaussians;  /* rows x 12 */
    double TracerDis;
    int wind_rows;     const double *wind;       /* rows x 2: lat_deg, u */
} sw2_params;

/* Build a world from parameters (the part of initialize() after parsing).
 * Returns 0 or a negative error (see sw2_last_error). */
int sw2_init_params(const sw2_params *P, sw *SW);

/* Parse a case file into parameters; tables are allocated and owned by the
 * library until sw2_free_params.  Returns 0 or a negative error. */
int sw2_read_case(const char *fname, sw2_params *P);
void sw2_free_params(sw2_params *P);

/* Evaluate one case-file <expr> (tinyexpr grammar of sppde_tinyexpr.c). */
double sw2_eval_expr(const char *expr, int *error);

/* Single-map halo update of one field (sppde_hupdate.c:274-363). */
void halo_update(double *x, map *M);

/* Host versions kept for post-processing (postproc/Postproc.c:166,189). */
void do_channel(sw *SW);
void calc_energy_balance(double *ke, double *ape, sw *SW);

/* SW2's per-step functions (sw.h:229-297), run on the GPU.  Each reads SW's
 * host arrays as the reference function does and writes back the cells it
 * writes: the interior of the output for the tendencies, the whole arrays
 * for the Laplacians, u or v in place for Coriolis.  One world at a time
 * (its device copy is cached between calls); not re-entrant; errors CRASH.
 * Results are SW2's bit for bit. */
void solve_tracer(double *Dtracer_n, double Dt, sw *SW);                  /* calc.c:241-275 */
void solve_cons_h(double *Deta_n, double Dt, sw *SW);                     /* calc.c:289-355 */
void dissipation_effect(double *Du_n, double *Dv_n, sw *SW);              /* calc.c:357-400 */
void do_eta_geofactor(double *Deta_nAB, double Dt, sw *SW);               /* calc.c:403-422 */
void solve_adv_u(double *Du_adv, double Dt, sw *SW);                      /* calc.c:424-460 */
void solve_adv_v(double *Dv_adv, double Dt, sw *SW);                      /* calc.c:462-498 */
void solve_pres_u(double *Du_pres, double Dt, sw *SW);                    /* calc.c:500-519 */
void solve_pres_v(double *Dv_pres, double Dt, sw *SW);                    /* calc.c:521-540 */
void solve_Coriolis_u(double *Du_nAB, double Dt, sw *SW);                 /* calc.c:542-569 */
void solve_Coriolis_v(double *Dv_nAB, double Dt, sw *SW);                 /* calc.c:571-598 */
void do_AdamsBashforth(double *Dvar_nAB, double *Dvar_n, double *Dvar_nM1, double *Dvar_nM2, int n,
                       map *M);                                           /* calc.c:651-699 */
double compute_Dt(double Courant, sw *SW);                                /* calc.c:701-738 */
void laplacian_vels(double *u_in, double *v_in, double *u_out, double *v_out, sw *SW); /* laplacian.c:51-111 */
void laplacian_eta(double *e_in, double *e_out, sw *SW);                  /* laplacian.c:114-151 */

/* Mass diagnostic (no counterpart in SW2): sum of H*Dxc*Dyc over the
 * interior in SW2's loop order. */
double sw2_mass(sw *SW);

/* Frame IO in SW2's MAGICSW2 format (io.c:27-171). */
void save_binary_sw(int frame, int nstep, double timedouble, char *bname, char *output_folder, sw *SW);
int load_binary_sw(int frame, int *nstep, double *timedouble, char *bname, char *binaries_folder, sw *SW);

/* Forms of the device step. */
#define SW2_FORM_FP64  0   /* plain FP64, same results as SW2's CPU path */
#define SW2_FORM_MIXED 1   /* anomalies about an FP64 mean state (FP32 value + FP32
                              * residual per cell): FP32 fluxes, limiters, pressure
                              * gradients and Laplacians; FP64 mean terms, Adams-Bashforth
                              * accumulation, Coriolis update and reductions.  Cases the
                              * fused step takes (sw2_session_fused): storms with or
                              * without tracer, vortices and del-6 (in a session;
                              * batches take cases without vortices or del-6). */
/* Flag or'ed into the form: run one kernel per phase of the step even where
 * the fused two-kernel step applies (same results; for cross-checks). */
#define SW2_PHASED 0x100

/* A device session owns the device copy of one world.  Creating it uploads
 * eta, u, v, tracer, perturbs, hB, windx, windy (halos included) and the
 * per-row metric tables; steps then run without host traffic except for the
 * events SW2 handles on the host (Gaussian injection, vortices). */
typedef struct sw2_session sw2_session;

int sw2_session_create(sw *SW, int form, sw2_session **out);
/* Advance nsteps steps.  n0 is the index of the step before the first one
 * (SW2's nstart), t the time before it; AB restarts at step n0+1 as SW2's
 * loop does (core.c:166 n-nstart). */
int sw2_session_begin(sw2_session *s, int nstart, double t);
int sw2_session_step(sw2_session *s, int nsteps);
/* Advance nsteps steps and return their device time in ms, measured with
 * HIP events on the session's stream. */
int sw2_session_time(sw2_session *s, int nsteps, double *ms);
/* Advance nsteps steps on the fused path and return the summed device time
 * of each of its two kernels, ms[0] the height kernel and ms[1] the momentum
 * kernel (HIP events around every launch on the session's stream; nk >= 2).
 * Fails on a session that steps one kernel per phase. */
int sw2_session_kernel_times(sw2_session *s, int nsteps, double *ms, int nk);
/* SW2's energy diagnostics (energy.c:25-55) and the mass sum H*Dxc*Dyc of
 * the current state, reduced on the device in FP64 (a fixed-order tree, not
 * SW2's serial order: for the FP64 form the host's calc_energy_balance and
 * sw2_mass give SW2's exact sums). */
int sw2_session_energy(sw2_session *s, double *ke, double *ape, double *mass);
/* No counterpart in SW2: sw2_session_energy for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops.  The same
 * reduction (same blocks, same order: the same three sums) of the session's
 * current state, written to d_out3 (caller-owned device memory, 3 doubles:
 * KE, APE, mass) on `stream` (a hipStream_t; NULL: the session's stream)
 * without host synchronisation and without allocating anything: its partial
 * sums live in a static device array of 3 x 1024 doubles (24 KiB) that every
 * GPU the library runs on holds until the process ends.  Any grid reduces to
 * those 1024 partials, so no size is refused.  Calls that overlap on one GPU
 * share that array and race; order them on one stream.  Call with the
 * session's device current. */
int sw2_session_energy_noalloc(sw2_session *s, double *d_out3, void *stream);
/* Copy eta, u, v, tracer, perturbs (halos included) back into SW's arrays.
 * A mixed-form session writes its FP64 mean plus its anomaly (value +
 * residual). */
int sw2_session_download(sw2_session *s, sw *SW);
/* First step whose CFL or stats guard failed (0 if none) and its message. */
int sw2_session_failed_step(sw2_session *s, char *msg, int msglen);
/* Dt_CFL of the most recent step (compute_Dt, calc.c:701-738). */
int sw2_session_last_cfl(sw2_session *s, double *dtcfl);
/* 1 when the session steps with the two fused kernels, 0 when it runs one
 * kernel per phase (grids narrower than 128 columns, a wind table that is
 * not one value per row, or SW2_PHASED). */
int sw2_session_fused(sw2_session *s);
/* SW2_FORM_FP64 or SW2_FORM_MIXED. */
int sw2_session_form(sw2_session *s);
/* Current step index and time. */
int sw2_session_state(sw2_session *s, int *n, double *t);
void sw2_session_destroy(sw2_session *s);

/* ---------------------------------------------------------------------------
 * Batches (no counterpart in SW2): many worlds of any sizes stepped together
 * in the mixed-precision form.  Every active member's tiles go through one
 * launch per phase (by default one workgroup per tile; persistent
 * workgroups as a set launch shape); the steps are captured in a hipGraph
 * and replayed, captured again when the tile count changes (members joining,
 * leaving, held or released) while launching one workgroup per tile.
 * Members with in-loop storms step in the replayed graph too: before each
 * replay the host evaluates their storms for the graph's steps and copies
 * the cells to the device, where the graph's storm kernels inject them.
 * Members must be cases the mixed form takes and share the limiter and
 * hyperviscosity order.
 * ------------------------------------------------------------------------- */
typedef struct sw2_batch sw2_batch;

/* Compute units of the current device (the unit of the launch shapes). */
int sw2_device_cus(void);
/* HBM bytes a member of this size occupies in the given form. */
size_t sw2_batch_footprint(const sw *SW, int form);
/* capacity: member slots; hbm_budget_bytes: admission limit on the sum of
 * member footprints (<= 0: none). */
int sw2_batch_create(int form, int capacity, double hbm_budget_bytes, sw2_batch **out);
/* Admit a world (its current state is the initial state) for nsteps steps.
 * Returns 0 and the member id, 1 when there is no room (slots or HBM
 * budget; nothing changed), or a negative error.  The world's upload
 * completes before it returns; it does not wait for steps the batch still
 * runs (the member joins at the next sw2_batch_step), so a server can admit
 * while a tick runs. */
int sw2_batch_add(sw2_batch *b, sw *SW, int nsteps, int *id);
/* Launch shape: workgroups of the height and momentum kernels and steps per
 * captured graph (the autotuner's knobs).  0 workgroups: one per tile of
 * the batch's members (the graph is captured again when members join or
 * leave); n > 0: n persistent workgroups striding over the tiles.  A new
 * batch starts with one workgroup per tile and 8 steps per graph. */
int sw2_batch_set_launch(sw2_batch *b, int blocks_eta, int blocks_uv, int steps_per_graph);
int sw2_batch_get_launch(sw2_batch *b, int *blocks_eta, int *blocks_uv, int *steps_per_graph);
/* Tile shape of the batch kernels, columns x rows: 32 x 16 (a new batch's),
 * 64 x 8 or 64 x 16 (one workgroup per tile only); any time between steps,
 * members included.  Results do not depend on the shape. */
int sw2_batch_set_tiles(sw2_batch *b, int tile_x, int tile_y);
int sw2_batch_get_tiles(sw2_batch *b, int *tile_x, int *tile_y);
/* Queue nsteps steps of every active member (each stops at its own last
 * step); returns without waiting. */
int sw2_batch_step(sw2_batch *b, int nsteps);
int sw2_batch_sync(sw2_batch *b);
/* Hold a member (paused != 0: its tiles leave the launches and its step
 * count stays) or release it, between steps; a scheduler's knob.  With one
 * workgroup per tile a change re-captures the step graph at the next step. */
int sw2_batch_pause(sw2_batch *b, int id, int paused);
/* Steps completed, whether the member's run ended, first failed step. */
int sw2_batch_member(sw2_batch *b, int id, int *steps_done, int *done, int *failed_step);
/* Ids of members whose runs ended; returns how many (<= maxids). */
int sw2_batch_finished(sw2_batch *b, int *ids, int maxids);
/* A member's state into SW's arrays (FP64 mean plus FP32 anomaly). */
int sw2_batch_download(sw2_batch *b, int id, sw *SW);
/* A member's energy and mass, reduced on the device in FP64. */
int sw2_batch_energy(sw2_batch *b, int id, double *ke, double *ape, double *mass);
/* Retire a member and free its slot and device memory. */
int sw2_batch_remove(sw2_batch *b, int id);
/* Members admitted and their HBM bytes. */
int sw2_batch_count(sw2_batch *b, size_t *used_bytes);
void sw2_batch_destroy(sw2_batch *b);

/* Last error message of this thread's most recent failing sw2_* call. */
const char *sw2_last_error(void);
/* Library build string (architecture, flags). */
const char *sw2_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* SW2_HIP_H */
