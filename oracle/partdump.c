/*
 * Test infrastructure: runs the reference's rank partition (nurbs_init,
 * mesh/nurbs/nurbs_init.c:9-199, compiled from the reference tree by
 * `make -C oracle part`) on an open uniform p = 2 mesh of nx x ny x nz
 * elements under mpiexec, and writes every rank's Nurbs fields as one JSON
 * object to <outdir>/rank_<r>.json.  Used only by tests/golden/make_partition.py
 * to write golden fixtures.
 *
 *   mpiexec -n P oracle/_ref/partdump nx ny nz outdir
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>

#include "nurbs.h"

static void ints(FILE *f, const char *name, const int *a, int n, int last)
{
    fprintf(f, "\"%s\": [", name);
    for (int i = 0; i < n; i++) fprintf(f, "%s%d", i ? ", " : "", a[i]);
    fprintf(f, "]%s\n", last ? "" : ",");
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    if (argc < 5) {
        fprintf(stderr, "usage: partdump nx ny nz outdir\n");
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    Nurbs nb;
    nb.porder = 2;
    nb.bc_periodic = (int *)calloc(3, sizeof(int));
    nb.nelem_global = (int *)malloc(3 * sizeof(int));
    nb.knotVector_global = (double **)malloc(3 * sizeof(double *));
    for (int d = 0; d < 3; d++) {
        int n = atoi(argv[1 + d]);
        nb.nelem_global[d] = n;
        double *k = (double *)malloc((size_t)(n + 5) * sizeof(double));
        /* open uniform knots on [0, 1] as nurbs_genuniform.c:34-38 computes them */
        const double l = 1.0 / n;
        for (int i = 0; i < n + 5; i++) k[i] = i < 3 ? 0.0 : (i >= n + 2 ? 1.0 : (i - 2) * l);
        nb.knotVector_global[d] = k;
    }
    nurbs_init(&nb);
    int rank, nproc;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &nproc);
    char path[4096];
    snprintf(path, sizeof(path), "%s/rank_%d.json", argv[4], rank);
    FILE *f = fopen(path, "w");
    if (!f) MPI_Abort(MPI_COMM_WORLD, 3);
    fprintf(f, "{\"rank\": %d, \"nproc\": %d,\n", rank, nproc);
    ints(f, "nelem", nb.nelem, 3, 0);
    ints(f, "nbasis", nb.nbasis, 3, 0);
    ints(f, "nknot", nb.nknot, 3, 0);
    for (int d = 0; d < 3; d++) {
        fprintf(f, "\"knot%d\": [", d);
        for (int i = 0; i < nb.nknot[d]; i++) fprintf(f, "%s%.17g", i ? ", " : "", nb.knotVector[d][i]);
        fprintf(f, "],\n");
    }
    ints(f, "boun", nb.boun, nb.nboun, 0);
    ints(f, "sendPart", nb.sendPart, nb.nsendPart, 0);
    ints(f, "sendBlock", nb.sendBlock, nb.nsendPart, 0);
    ints(f, "recvPart", nb.recvPart, nb.nrecvPart, 0);
    ints(f, "recvBlock", nb.recvBlock, nb.nrecvPart, 0);
    ints(f, "universalIdx", nb.universalIdx_temp, nb.nbasis[0] * nb.nbasis[1] * nb.nbasis[2], 1);
    fprintf(f, "}\n");
    fclose(f);
    MPI_Finalize();
    return 0;
}
