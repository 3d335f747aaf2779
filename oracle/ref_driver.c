/*
 * ref_driver.c -- flat C entry points over the reference's own mapmat code
 * (algebra/src/mapmat/{mapmat,csort,cindex,...}.c compiled from source by
 * oracle/Makefile into oracle/_ref/libref_mapmat.so).  TEST INFRASTRUCTURE:
 * used only to pin the oracle's index tables and pointing products.
 *
 * MPI is initialised in singleton mode on first use; MatInit/MatComShape run
 * on MPI_COMM_WORLD of size 1, where every reduction is an identity.
 */
#include <mpi.h>
#include <stdlib.h>
#include <string.h>

#include "mapmat.h"

static void ensure_mpi(void) {
    int fl = 0;
    MPI_Initialized(&fl);
    if (!fl) MPI_Init(NULL, NULL);
}

/* MatInit (mapmat.c:54-73) with flag 6 (ALLREDUCE, the TOAST default):
 * indices are rewritten in place; lindices copied out. */
int ref_local_shape(int m, int nnz, int *indices, double *values, int *lindices, int *lcount,
                    int *trash_pix) {
    ensure_mpi();
    Mat A;
    memset(&A, 0, sizeof(A));
    A.trash_pix = 0;
    MatInit(&A, m, nnz, indices, values, 6, MPI_COMM_WORLD);
    memcpy(lindices, A.lindices, (size_t) A.lcount * sizeof(int));
    *lcount = A.lcount;
    *trash_pix = A.trash_pix;
    MatFree(&A);
    return 0;
}

/* MatVecProd / TrMatVecProd (mapmat.c:704-729, 819-859) on global indices:
 * builds the Mat from a copy of the indices, then applies the product. */
int ref_products(int m, int nnz, const int *gindices, double *values, const double *x_in,
                 double *y_out, const double *y_in, double *x_out) {
    ensure_mpi();
    Mat A;
    memset(&A, 0, sizeof(A));
    int *idx = (int *) malloc((size_t) m * nnz * sizeof(int));
    memcpy(idx, gindices, (size_t) m * nnz * sizeof(int));
    A.trash_pix = 0;
    MatInit(&A, m, nnz, idx, values, 6, MPI_COMM_WORLD);
    if (x_in && y_out) MatVecProd(&A, (double *) x_in, y_out, 0);
    if (y_in && x_out) TrMatVecProd(&A, (double *) y_in, x_out, 0);
    MatFree(&A);
    free(idx);
    return 0;
}
