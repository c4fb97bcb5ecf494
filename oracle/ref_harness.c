/*
 * ref_harness.c — thin C entry points around the reference's own Faddeeva
 * and SVD routines, for writing golden vectors in the build container.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
 * sources where they lie (Src/FADDEEVA.c, Src/SVD.c, Src/ALLOCATION.c,
 * Src/LOG_ERROR.c) into oracle/_ref/libref.so, which stays out of git.
 * The rest of the reference hot path (Src/ME_SOLVER.c, Src/LM_FIT.c,
 * Src/INV_INIT.c) includes fitsio.h, which this image lacks, so it is not
 * built.
 */
#include <stdint.h>
#include <string.h>
#include "FADDEEVA.h"
#include "SVD.h"

void ref_faddeeva_vec(int64_t n, const double *nu, const double *y, int nfigures,
                      double *H, double *L)
{
    STRUCT_FADDEEVA F;
    memset(&F, 0, sizeof(F));
    F.nfigures = nfigures;
    Faddeeva_init(&F);
    for (int64_t i = 0; i < n; i++) Faddeeva(nu[i], y[i], &H[i], &L[i], &F);
    free(F.Expa2n2);
}

/* a [n][n] in, replaced by U; w [n]; v [n][n] */
void ref_svd(double *a, int n, double *w, double *v)
{
    STRUCT_MATRIX A = MATRIX(0, n - 1, 0, n - 1, enum_dbl, false);
    STRUCT_MATRIX V = MATRIX(0, n - 1, 0, n - 1, enum_dbl, false);
    memcpy(A.data, a, sizeof(double) * n * n);
    svd_dbl(&A, w, &V);
    memcpy(a, A.data, sizeof(double) * n * n);
    memcpy(v, V.data, sizeof(double) * n * n);
    FREE_MATRIX(A);
    FREE_MATRIX(V);
}

void ref_svbksb(const double *u, const double *v, const double *w, int n, const double *b, double *x)
{
    STRUCT_MATRIX U = MATRIX(0, n - 1, 0, n - 1, enum_dbl, false);
    STRUCT_MATRIX V = MATRIX(0, n - 1, 0, n - 1, enum_dbl, false);
    memcpy(U.data, u, sizeof(double) * n * n);
    memcpy(V.data, v, sizeof(double) * n * n);
    svbksb_double(&U, &V, w, b, x);
    FREE_MATRIX(U);
    FREE_MATRIX(V);
}
