/*
 * ref_main.c -- SW2's run (main/Main.c:22-91) on a case file without m4
 * constructs, for timing the reference as the CPU baseline:
 *
 *     mpiexec -n P oracle/_ref/sw2ref npx npy case.sw bname outdir
 *
 * TEST INFRASTRUCTURE ONLY (see ref_harness.c).  Prints SW2's own timer
 * table (cr_info, sppde/sppde_cr.c:131-163), whose time_step00 line holds
 * the mean wall time per step, and the wall time of loop() on rank 0.
 */
#include <stdio.h>
#include <stdlib.h>

#include "mpi.h"
#include "sppde.h"
#include "sw.h"

int ref_start(const char *outdir);
void ref_stop(void);
sw *ref_world(int npx, int npy, const char *casefile);
void ref_free(sw *SW);

int main(int argc, char **argv) {
    checkr(MPI_Init(&argc, &argv), "init");
    if (argc != 6) CRASH("run as: sw2ref npx npy case.sw bname output_folder");
    int npx = atoi(argv[1]), npy = atoi(argv[2]);
    ref_start(argv[5]);
    cr_start("TOTAL", 0);
    sw *SW = ref_world(npx, npy, argv[3]);
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    loop(argv[4], argv[5], SW);
    MPI_Barrier(MPI_COMM_WORLD);
    double t1 = MPI_Wtime();
    ref_free(SW);
    cr_end("TOTAL", 0);
    pprintf("REF loop_wall_s %.6f ranks %d npx %d npy %d\n", t1 - t0, quants(), npx, npy);
    cr_info();
    ref_stop();
    return 0;
}
