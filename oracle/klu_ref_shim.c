/* Thin C shim over the reference's vendored SuiteSparse KLU (vendor/suitesparse, v7.12.2).
 *
 * TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/libklu_ref.so together with the KLU/AMD/BTF
 * sources compiled where they lie in the reference tree (see oracle/Makefile).  It exposes the
 * symbolic ordering (BTF P/Q/R + per-block AMD), the numeric pivot order (Pnum) and the L/U factors
 * that the reference's KluNative backend is differentially tested against
 * (src/klu_native/ffi_oracle.rs:27-56, src/sparse_klu.rs:105-200).
 *
 * KluNative runs unscaled (src/klu_native/mod.rs:44-58), so callers pass scale = 0 to compare
 * against it; scale = 2 is KLU's own default (klu_defaults.c).
 */
#include <stdlib.h>
#include <string.h>
#include "klu.h"

typedef struct {
    klu_common common;
    klu_symbolic *sym;
    klu_numeric *num;
    int n;
} kref_t;

void *kref_create(int n, const int *Ap, const int *Ai, const double *Ax, int scale)
{
    kref_t *h = (kref_t *) calloc(1, sizeof(kref_t));
    if (!h) return NULL;
    klu_defaults(&h->common);
    h->common.scale = scale;
    h->n = n;
    h->sym = klu_analyze(n, (int32_t *) Ap, (int32_t *) Ai, &h->common);
    if (!h->sym) { free(h); return NULL; }
    h->num = klu_factor((int32_t *) Ap, (int32_t *) Ai, (double *) Ax, h->sym, &h->common);
    if (!h->num) {
        klu_free_symbolic(&h->sym, &h->common);
        free(h);
        return NULL;
    }
    return h;
}

void kref_destroy(void *p)
{
    kref_t *h = (kref_t *) p;
    if (!h) return;
    klu_free_numeric(&h->num, &h->common);
    klu_free_symbolic(&h->sym, &h->common);
    free(h);
}

/* counts[0..3] = nblocks, lnz (incl. diagonal), unz (incl. diagonal), nzoff */
void kref_counts(void *p, int *counts)
{
    kref_t *h = (kref_t *) p;
    counts[0] = h->num->nblocks;
    counts[1] = h->num->lnz;
    counts[2] = h->num->unz;
    counts[3] = h->num->nzoff;
}

/* Symbolic orders before numeric pivoting: P, Q (size n) and R (size nblocks+1). */
void kref_symbolic(void *p, int *P, int *Q, int *R)
{
    kref_t *h = (kref_t *) p;
    memcpy(P, h->sym->P, sizeof(int) * h->n);
    memcpy(Q, h->sym->Q, sizeof(int) * h->n);
    memcpy(R, h->sym->R, sizeof(int) * (h->sym->nblocks + 1));
}

/* Full factorization in KLU's extracted form: L unit lower (diagonal stored), U upper (diagonal
 * stored), F off-diagonal BTF blocks, final row order P (= Pnum), column order Q, blocks R. */
int kref_extract(void *p, int *Lp, int *Li, double *Lx, int *Up, int *Ui, double *Ux, int *Fp, int *Fi,
                 double *Fx, int *P, int *Q, int *R)
{
    kref_t *h = (kref_t *) p;
    return klu_extract(h->num, h->sym, Lp, Li, Lx, Up, Ui, Ux, Fp, Fi, Fx, P, Q, NULL, R, &h->common);
}

/* Numeric refactor with the pivot order kept, then solve A x = b in place.  Returns 1 on success. */
int kref_refactor_solve(void *p, const int *Ap, const int *Ai, const double *Ax, double *b)
{
    kref_t *h = (kref_t *) p;
    if (!klu_refactor((int32_t *) Ap, (int32_t *) Ai, (double *) Ax, h->sym, h->num, &h->common)) return 0;
    return klu_solve(h->sym, h->num, h->n, 1, b, &h->common);
}

/* Solve with the current factorization. */
int kref_solve(void *p, double *b)
{
    kref_t *h = (kref_t *) p;
    return klu_solve(h->sym, h->num, h->n, 1, b, &h->common);
}
