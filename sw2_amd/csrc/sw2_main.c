/*
 * SW -- command-line driver of the GPU library, the same program as SW2's
 * main/Main.c:25-90 built without MPI (one process owns the whole domain on
 * one GPU).  Usage and output are SW2's:
 *
 *     SW npx npy input_file output_folder
 *
 * input_file is the case name without ".sw"; npx*npy must be 1.
 */
#include <stdio.h>
#include <string.h>

#include "sw2_hip.h"

int main(int argc, char **argv) {
    char output_folder[MAXOUTLEN];
    char bname[MAXOUTLEN];

    if (argc != 5)
        CRASH("SW VERSION=%s COMPILED %s %s\nRun as\nSW npx npy input_file output_folder", VERSION, __DATE__, __TIME__);
    if (strlen(argv[4]) >= MAXOUTLEN) CRASH("Output path=<%s> is too long", argv[4]);
    if (strlen(argv[3]) >= MAXOUTLEN) CRASH("Input name=<%s> is too long", argv[3]);
    snprintf(output_folder, sizeof output_folder, "%s", argv[4]);
    snprintf(bname, sizeof bname, "%s", argv[3]);

    init_pprintf(2, 0, output_folder);
    cr_reset();
    pprintf("SW STARTED VERSION: %s\n", VERSION);
    pprintf("COMPILED: %s %s\n", __DATE__, __TIME__);
    pprintf("%d processes started\n", quants());
    cr_start("TOTAL", 0);

    sw myworld;
    sw *SW = &myworld;
    int npx = 0, npy = 0;
    sscanf(argv[1], "%d", &npx);
    sscanf(argv[2], "%d", &npy);
    pprintf("npx=%d npy=%d \n", npx, npy);

    initialize(npx, npy, 0, bname, 1, SW);
    loop(bname, output_folder, SW);
    destroy(SW);

    cr_end("TOTAL", 0);
    pprintf("PROGRAM ENDED\n");
    cr_info();
    end_pprintf();
    return 0;
}
