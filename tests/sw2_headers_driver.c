/*
 * sw2_headers_driver.c -- a caller compiled against SW2's own headers
 * (sw.h, sppde/sppde.h) and linked against libsw2.so, as SW2's main/Main.c or
 * postproc/Postproc.c would be.  It reads every value through SW2's struct
 * definitions, so it checks that the library's map and sw have SW2's layout.
 *
 *     sw2_headers_driver bname outdir nsteps raw_out
 *
 * initialize() the case <bname>.sw; print the sizes, the map and the first
 * rows of the geometry; do_channel() and calc_energy_balance() (the host
 * functions Postproc.c calls); then, with nsteps > 0, loop() over the first
 * nsteps steps on the GPU and write eta, u, v (whole arrays, M->S doubles
 * each) to raw_out; destroy().  Built by the tests where the reference
 * sources are present; test infrastructure only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "sppde.h"
#include "sw.h"

int main(int argc, char **argv) {
    if (argc != 5) {
        fprintf(stderr, "usage: sw2_headers_driver bname outdir nsteps raw_out\n");
        return 2;
    }
    int nsteps = atoi(argv[3]);
    init_pprintf(2, 0, argv[2]);
    cr_reset();
    sw world;
    sw *SW = &world;
    initialize(1, 1, 0, argv[1], 0, SW);
    printf("LAYOUT nx %d ny %d N %d S %d sh %d l0 %d %d l1 %d %d s %d %d per %d %d bc %d %d %d %d nb %d %d %d %d %d %d\n",
           SW->nx, SW->ny, SW->N, SW->M->S, SW->M->sh, SW->M->l0[1], SW->M->l0[2], SW->M->l1[1], SW->M->l1[2],
           SW->M->s[1], SW->M->s[2], SW->M->per[1], SW->M->per[2], SW->M->bc0[1], SW->M->bc1[1], SW->M->bc0[2],
           SW->M->bc1[2], SW->M->nb_e, SW->M->nb_w, SW->M->nb_n, SW->M->nb_s, SW->M->nb_t, SW->M->nb_b);
    printf("VALUES depth %.17g Omega %.17g gE %.17g nu4 %.17g Dt %.17g sigma %.17g gauss %d vort %d\n", SW->depth,
           SW->Omega, SW->gE, SW->nu4, SW->Dt, SW->sigma_dis, SW->num_Gaussians, SW->num_vortices);
    printf("ROWS Dxc %.17g %.17g gc %.17g fc %.17g windx %.17g\n", ac(SW->Dxc, 1, 1, SW->M), ac(SW->Dxc, 1, SW->ny, SW->M),
           ac(SW->gc, 1, 1, SW->M), ac(SW->fc, 1, 1, SW->M), ac(SW->windx, 1, 1, SW->M));
    do_channel(SW);
    double ke = 0, ape = 0;
    calc_energy_balance(&ke, &ape, SW);
    printf("ENERGY0 %.17g %.17g\n", ke, ape);
    if (nsteps > 0) {
        SW->N = nsteps;
        SW->t1 = SW->t0 + nsteps * SW->Dt;
        loop(argv[1], argv[2], SW);
        calc_energy_balance(&ke, &ape, SW);
        printf("ENERGY1 %.17g %.17g\n", ke, ape);
        FILE *f = fopen(argv[4], "wb");
        if (!f) return 3;
        fwrite(SW->eta, sizeof(double), SW->M->S, f);
        fwrite(SW->u, sizeof(double), SW->M->S, f);
        fwrite(SW->v, sizeof(double), SW->M->S, f);
        fclose(f);
    }
    destroy(SW);
    end_pprintf();
    return 0;
}
