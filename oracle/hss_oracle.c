/*
 * hss_oracle.c -- CPU oracle for the HSS matvec and the ULV solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithms in
 * plain C so that the GPU path can be checked against them; nothing in the
 * product (libhss_hip.so, the Python package) links, loads or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * It works on the same packed input as hss_mat_create (include/hss_hip.h):
 * a breadth-first node table and a column-major generator blob.
 *
 * Restated functions (reference = bonevbs/HssMatrices.jl 0.1.6):
 *   hsso_matmul        src/matmul.jl:18-28   mul!(C, hssA, B, alpha, beta)
 *     mm_up            src/matmul.jl:32-42   _matmatup
 *     mm_down          src/matmul.jl:44-62   _matmatdown!
 *   hsso_ulvfactsolve  src/ulvfactor.jl:10-19  ulvfactsolve
 *     ulv_reduce       src/ulvfactor.jl:22-57  _ulvreduce!  (4-argument form)
 *     ulv_rec          src/ulvfactor.jl:60-96  _ulvfactsolve!
 *     ulv_topdown      src/ulvfactor.jl:98-107 _ulvfactsolve_topdown!
 *
 * The arithmetic goes through the same LAPACK/BLAS routines the reference
 * calls through Julia's LinearAlgebra (dgemm, dgeqlf, dormql, dgelqf, dormlq,
 * dtrsm, dgetrf, dgetrs).  The caller passes their addresses in hsso_lapack
 * (the Python driver takes them from scipy's cython_lapack/cython_blas, i.e.
 * OpenBLAS), so this file has no link-time dependency.
 *
 * Dense "\" at the root (ulvfactor.jl:83) and for a leaf-only matrix
 * (ulvfactor.jl:11-12) is restated as LU with partial pivoting
 * (dgetrf + dgetrs), which is what Julia's "\" does for a general square
 * matrix; an exactly zero pivot returns HSSO_SINGULAR (Julia raises
 * SingularException there).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef void (*dgemm_t)(char*, char*, int*, int*, int*, double*, double*, int*, double*, int*, double*,
                        double*, int*);
typedef void (*dgeqlf_t)(int*, int*, double*, int*, double*, double*, int*, int*);
typedef void (*dormql_t)(char*, char*, int*, int*, int*, double*, int*, double*, double*, int*, double*,
                         int*, int*);
typedef void (*dtrsm_t)(char*, char*, char*, char*, int*, int*, double*, double*, int*, double*, int*);
typedef void (*dgetrf_t)(int*, int*, double*, int*, int*, int*);
typedef void (*dgetrs_t)(char*, int*, int*, double*, int*, int*, double*, int*, int*);

typedef struct hsso_lapack {
    dgemm_t dgemm;
    dgeqlf_t dgeqlf;
    dormql_t dormql;
    dgeqlf_t dgelqf; /* same signature as dgeqlf */
    dormql_t dormlq; /* same signature as dormql */
    dtrsm_t dtrsm;
    dgetrf_t dgetrf;
    dgetrs_t dgetrs;
} hsso_lapack;

/* node record, identical to hss_node_desc in include/hss_hip.h */
typedef struct hsso_node {
    int32_t parent, left, right, depth, height, m, n, ku, kv, pad;
    int64_t row_off, col_off, off_D, off_U, off_V, off_B12, off_B21, off_R, off_W;
} hsso_node;

enum { HSSO_OK = 0, HSSO_DIM = 1, HSSO_SINGULAR = 3, HSSO_OOM = 5, HSSO_LAPACK = 9 };

/* ------------------------------------------------------------------------ */
/* small column-major matrix helpers                                         */
/* ------------------------------------------------------------------------ */
typedef struct { double* a; int r, c; } mat;

static mat mnew(int r, int c) {
    mat x; x.r = r; x.c = c;
    size_t n = (size_t)(r > 0 ? r : 0) * (size_t)(c > 0 ? c : 0);
    x.a = (double*)calloc(n ? n : 1, sizeof(double));
    return x;
}
static void mfree(mat* x) { free(x->a); x->a = NULL; x->r = x->c = 0; }
static mat mcopy(const double* src, int r, int c, int ld) {
    mat x = mnew(r, c);
    for (int j = 0; j < c; ++j) memcpy(x.a + (size_t)j * r, src + (size_t)j * ld, sizeof(double) * (size_t)r);
    return x;
}
/* x[r0:r1, c0:c1] as a new matrix */
static mat msub(const mat* x, int r0, int r1, int c0, int c1) {
    return mcopy(x->a + (size_t)c0 * x->r + r0, r1 - r0, c1 - c0, x->r > 0 ? x->r : 1);
}
static mat mview_blob(const double* blob, int64_t off, int r, int c) {
    /* copy of a generator block (the reference copies leaf generators, ulvfactor.jl:63) */
    if (r * c == 0 || off < 0) return mnew(r, c);
    return mcopy(blob + off, r, c, r);
}

/* C = alpha*op(A)*op(B) + beta*C with BLAS dgemm; handles empty dims */
static void gemm(const hsso_lapack* L, char ta, char tb, int m, int n, int k, double alpha, const double* A,
                 int lda, const double* B, int ldb, double beta, double* C, int ldc) {
    if (m <= 0 || n <= 0) return;
    if (k <= 0) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) C[i + (size_t)j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + (size_t)j * ldc];
        return;
    }
    if (lda < 1) lda = 1;
    if (ldb < 1) ldb = 1;
    if (ldc < 1) ldc = 1;
    L->dgemm(&ta, &tb, &m, &n, &k, &alpha, (double*)A, &lda, (double*)B, &ldb, &beta, C, &ldc);
}
/* ------------------------------------------------------------------------ */
/* matvec: reference src/matmul.jl:18-62                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const hsso_lapack* L;
    const hsso_node* nd;
    const double* blob;
    int nrhs;
    mat* Z; /* per node up-sweep result (BinaryNode Z tree of matmul.jl:34,39) */
} mmctx;

/* matmul.jl:32-42 ; B points at the node's first row of the RHS block */
static void mm_up(mmctx* c, int i, const double* B, int ldb) {
    const hsso_node* x = &c->nd[i];
    if (x->left < 0) {
        /* Z = V' * B   (matmul.jl:34) */
        c->Z[i] = mnew(x->kv, c->nrhs);
        gemm(c->L, 'T', 'N', x->kv, c->nrhs, x->n, 1.0, c->blob + (x->off_V < 0 ? 0 : x->off_V), x->n, B, ldb,
             0.0, c->Z[i].a, x->kv);
        return;
    }
    const hsso_node* l = &c->nd[x->left];
    const hsso_node* r = &c->nd[x->right];
    mm_up(c, x->left, B, ldb);
    mm_up(c, x->right, B + l->n, ldb);
    /* Z = W1'*Z1 .+ W2'*Z2  (matmul.jl:39); W1/W2 are stored with the children */
    c->Z[i] = mnew(x->kv, c->nrhs);
    mat t1 = mnew(x->kv, c->nrhs), t2 = mnew(x->kv, c->nrhs);
    gemm(c->L, 'T', 'N', x->kv, c->nrhs, l->kv, 1.0, c->blob + (l->off_W < 0 ? 0 : l->off_W), l->kv,
         c->Z[x->left].a, l->kv, 0.0, t1.a, x->kv);
    gemm(c->L, 'T', 'N', x->kv, c->nrhs, r->kv, 1.0, c->blob + (r->off_W < 0 ? 0 : r->off_W), r->kv,
         c->Z[x->right].a, r->kv, 0.0, t2.a, x->kv);
    for (size_t e = 0; e < (size_t)x->kv * c->nrhs; ++e) c->Z[i].a[e] = t1.a[e] + t2.a[e];
    mfree(&t1); mfree(&t2);
}

/* matmul.jl:44-62 ; F == NULL at the root */
static void mm_down(mmctx* c, int i, double* C, int ldc, const double* B, int ldb, const mat* F, double alpha,
                    double beta) {
    const hsso_node* x = &c->nd[i];
    if (x->left < 0) {
        /* mul!(C, D, B, alpha, beta); mul!(C, U, F, alpha, 1)  (matmul.jl:46-47) */
        gemm(c->L, 'N', 'N', x->m, c->nrhs, x->n, alpha, c->blob + x->off_D, x->m, B, ldb, beta, C, ldc);
        if (F) gemm(c->L, 'N', 'N', x->m, c->nrhs, x->ku, alpha, c->blob + (x->off_U < 0 ? 0 : x->off_U), x->m,
                    F->a, F->r, 1.0, C, ldc);
        return;
    }
    const hsso_node* l = &c->nd[x->left];
    const hsso_node* r = &c->nd[x->right];
    const mat* Z1 = &c->Z[x->left];
    const mat* Z2 = &c->Z[x->right];
    /* F1 = B12*Z2 + R1*F ; F2 = B21*Z1 + R2*F  (matmul.jl:51-57) */
    mat F1 = mnew(l->ku, c->nrhs), F2 = mnew(r->ku, c->nrhs);
    gemm(c->L, 'N', 'N', l->ku, c->nrhs, r->kv, 1.0, c->blob + (x->off_B12 < 0 ? 0 : x->off_B12), l->ku, Z2->a,
         r->kv, 0.0, F1.a, l->ku);
    gemm(c->L, 'N', 'N', r->ku, c->nrhs, l->kv, 1.0, c->blob + (x->off_B21 < 0 ? 0 : x->off_B21), r->ku, Z1->a,
         l->kv, 0.0, F2.a, r->ku);
    if (F) {
        mat t1 = mnew(l->ku, c->nrhs), t2 = mnew(r->ku, c->nrhs);
        gemm(c->L, 'N', 'N', l->ku, c->nrhs, x->ku, 1.0, c->blob + (l->off_R < 0 ? 0 : l->off_R), l->ku, F->a,
             F->r, 0.0, t1.a, l->ku);
        gemm(c->L, 'N', 'N', r->ku, c->nrhs, x->ku, 1.0, c->blob + (r->off_R < 0 ? 0 : r->off_R), r->ku, F->a,
             F->r, 0.0, t2.a, r->ku);
        for (size_t e = 0; e < (size_t)l->ku * c->nrhs; ++e) F1.a[e] += t1.a[e];
        for (size_t e = 0; e < (size_t)r->ku * c->nrhs; ++e) F2.a[e] += t2.a[e];
        mfree(&t1); mfree(&t2);
    }
    mm_down(c, x->left, C, ldc, B, ldb, &F1, alpha, beta);
    mm_down(c, x->right, C + l->m, ldc, B + l->n, ldb, &F2, alpha, beta);
    mfree(&F1); mfree(&F2);
}

/* mul!(C, hssA, B, alpha, beta), matmul.jl:18-28.  C is m x nrhs, B is n x nrhs. */
int hsso_matmul(const hsso_lapack* L, const hsso_node* nodes, int nn, const double* blob, int nrhs, double alpha,
                const double* B, int ldb, double beta, double* C, int ldc) {
    if (nn <= 0 || nrhs < 0) return HSSO_DIM;
    const hsso_node* root = &nodes[0];
    if (nrhs == 0) return HSSO_OK;
    if (root->left < 0) {
        gemm(L, 'N', 'N', root->m, nrhs, root->n, alpha, blob + root->off_D, root->m, B, ldb, beta, C, ldc);
        return HSSO_OK;
    }
    mmctx c = {L, nodes, blob, nrhs, (mat*)calloc((size_t)nn, sizeof(mat))};
    if (!c.Z) return HSSO_OOM;
    mm_up(&c, 0, B, ldb);
    mm_down(&c, 0, C, ldc, B, ldb, NULL, alpha, beta);
    for (int i = 0; i < nn; ++i) mfree(&c.Z[i]);
    free(c.Z);
    return HSSO_OK;
}

/* ------------------------------------------------------------------------ */
/* ULV factor-solve: reference src/ulvfactor.jl:10-107                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int* cols;   /* length ncols: global unknown indices of this node's columns */
    int ncols;
    mat lq;      /* LQ factor of the node (rows = reflectors) */
    double* tau; /* ntau reflector scalars */
    int ntau;
    int has;     /* 0 at the root (ulvfactor.jl:85) */
} qvnode;

typedef struct {
    const hsso_lapack* L;
    const hsso_node* nd;
    const double* blob;
    int nrhs;
    double* z; /* n x nrhs solution, ld = ldz */
    int ldz;
    qvnode* QV;
    int info;
} ulvctx;

static int lapack_lwork(double w) { int l = (int)w; return l > 1 ? l : 1; }

/* apply Q or Q^T of a QL/LQ factorization; f = dormql or dormlq */
static int orm(dormql_t f, char side, char trans, int m, int n, int k, double* A, int lda, double* tau, double* C,
               int ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    int info = 0, lwork = -1;
    double wq = 0.0;
    if (lda < 1) lda = 1;
    if (ldc < 1) ldc = 1;
    f(&side, &trans, &m, &n, &k, A, &lda, tau, C, &ldc, &wq, &lwork, &info);
    lwork = lapack_lwork(wq);
    double* work = (double*)malloc(sizeof(double) * (size_t)lwork);
    f(&side, &trans, &m, &n, &k, A, &lda, tau, C, &ldc, work, &lwork, &info);
    free(work);
    return info;
}
static int fact(dgeqlf_t f, int m, int n, double* A, int lda, double* tau) {
    if (m <= 0 || n <= 0) return 0;
    int info = 0, lwork = -1;
    double wq = 0.0;
    f(&m, &n, A, &lda, tau, &wq, &lwork, &info);
    lwork = lapack_lwork(wq);
    double* work = (double*)malloc(sizeof(double) * (size_t)lwork);
    f(&m, &n, A, &lda, tau, work, &lwork, &info);
    free(work);
    return info;
}

/* Result of one node (return tuple of _ulvfactsolve!, ulvfactor.jl:95). */
typedef struct {
    mat b, u, D, U, V;
    int* cols;
    int ncols;
    int nk;
} ulvres;

/* _ulvreduce!(D, U, V, b), ulvfactor.jl:22-57.  Consumes D, U, V, b and
 * returns the reduced ones; zloc/u/mk/nk/lq through the out parameters. */
static void ulv_reduce(ulvctx* c, mat* D, mat* U, mat* V, mat* b, mat* zloc, mat* u, int* mk, int* nkout,
                       mat* lq, double** tau, int* ntau) {
    const hsso_lapack* L = c->L;
    int m = D->r, n = D->c;
    int k = U->c < m ? U->c : m;           /* ulvfactor.jl:27 */
    int nk = (m - k) < n ? (m - k) : n;    /* ulvfactor.jl:28 */
    int nrhs = b->c;
    if (k >= m) {                          /* ulvfactor.jl:32-36 */
        *u = mnew(V->c, nrhs);
        *zloc = mnew(0, nrhs);
        *lq = mnew(0, n);
        *tau = NULL; *ntau = 0;
        *mk = m - k; *nkout = 0;
        return;
    }
    /* QL of the row generators, apply to D and b  (ulvfactor.jl:39-42) */
    double* tql = (double*)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
    int info = fact(L->dgeqlf, m, k, U->a, m, tql);
    if (info) c->info = HSSO_LAPACK;
    mat Unew = mnew(k, k);                 /* U = tril(U[end-k+1:end, :])  (:40) */
    for (int j = 0; j < k; ++j)
        for (int i = j; i < k; ++i) Unew.a[i + (size_t)j * k] = U->a[(m - k + i) + (size_t)j * m];
    if (orm(L->dormql, 'L', 'T', m, n, k, U->a, m, tql, D->a, m)) c->info = HSSO_LAPACK;
    if (orm(L->dormql, 'L', 'T', m, nrhs, k, U->a, m, tql, b->a, m)) c->info = HSSO_LAPACK;
    free(tql);
    mfree(U);
    *U = Unew;
    /* LQ of the first m-k rows of D  (:44) */
    mat lqf = msub(D, 0, m - k, 0, n);
    int ntq = (m - k) < n ? (m - k) : n;
    double* tlq = (double*)calloc((size_t)(ntq > 0 ? ntq : 1), sizeof(double));
    if (fact(L->dgelqf, m - k, n, lqf.a, m - k, tlq)) c->info = HSSO_LAPACK;
    /* L1 = tril(lqf)[:, 1:nk]  (:45-46) */
    mat L1 = mnew(m - k, nk);
    for (int j = 0; j < nk; ++j)
        for (int i = j; i < m - k; ++i) L1.a[i + (size_t)j * (m - k)] = lqf.a[i + (size_t)j * (m - k)];
    /* L2 = D[end-k+1:end, :] * Q'  (:47) */
    mat L2 = msub(D, m - k, m, 0, n);
    if (orm(L->dormlq, 'R', 'T', k, n, ntq, lqf.a, m - k, tlq, L2.a, k)) c->info = HSSO_LAPACK;
    /* zloc = L1 \ b[ind, :]  (:48) -- requires m-k == nk */
    mat zl = msub(b, 0, m - k, 0, nrhs);
    if (nk > 0 && nrhs > 0) {
        char s = 'L', ul = 'L', tr = 'N', dg = 'N';
        double one = 1.0;
        int mm = m - k, nn = nrhs, lda = m - k, ldb = m - k;
        L->dtrsm(&s, &ul, &tr, &dg, &mm, &nn, &one, L1.a, &lda, zl.a, &ldb);
    }
    /* b = b[cind, :] .- L2[:, 1:nk] * zloc  (:49) */
    mat bn = msub(b, m - k, m, 0, nrhs);
    {
        mat t = mnew(k, nrhs);
        gemm(L, 'N', 'N', k, nrhs, nk, 1.0, L2.a, k, zl.a, m - k, 0.0, t.a, k);
        for (size_t e = 0; e < (size_t)k * nrhs; ++e) bn.a[e] -= t.a[e];
        mfree(&t);
    }
    /* V = Q * V  (:50) ; u = V[ind, :]' * zloc  (:51) */
    if (orm(L->dormlq, 'L', 'N', n, V->c, ntq, lqf.a, m - k, tlq, V->a, n)) c->info = HSSO_LAPACK;
    *u = mnew(V->c, nrhs);
    gemm(L, 'T', 'N', V->c, nrhs, m - k, 1.0, V->a, n, zl.a, m - k, 0.0, u->a, V->c);
    /* D = L2[:, nk+1:end] ; V = V[nk+1:end, :]  (:53-54) */
    mat Dn = msub(&L2, 0, k, nk, n);
    mat Vn = msub(V, nk, n, 0, V->c);
    mfree(D); *D = Dn;
    mfree(V); *V = Vn;
    mfree(b); *b = bn;
    mfree(&L1); mfree(&L2);
    *zloc = zl;
    *lq = lqf;
    *tau = tlq; *ntau = ntq;
    *mk = m - k; *nkout = nk;
}

static void store_zloc(ulvctx* c, const int* cols, int mk, const mat* zloc) {
    for (int j = 0; j < c->nrhs; ++j)
        for (int i = 0; i < mk; ++i) c->z[cols[i] + (size_t)j * c->ldz] = zloc->a[i + (size_t)j * zloc->r];
}

/* _ulvfactsolve!(hssA, b, z, co; rootnode), ulvfactor.jl:60-96.  Consumes b. */
static ulvres ulv_rec(ulvctx* c, int i, mat b, int co, int rootnode) {
    const hsso_node* x = &c->nd[i];
    ulvres res;
    memset(&res, 0, sizeof(res));
    if (x->left < 0) {
        /* leaf (:61-66) */
        int n = x->n;
        int* cols = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int j = 0; j < n; ++j) cols[j] = co + j;
        mat D = mview_blob(c->blob, x->off_D, x->m, x->n);
        mat U = mview_blob(c->blob, x->off_U, x->m, x->ku);
        mat V = mview_blob(c->blob, x->off_V, x->n, x->kv);
        mat zloc, u, lq; double* tau; int ntau, mk, nk;
        ulv_reduce(c, &D, &U, &V, &b, &zloc, &u, &mk, &nk, &lq, &tau, &ntau);
        store_zloc(c, cols, mk, &zloc);
        mfree(&zloc);
        qvnode* q = &c->QV[i];
        q->ncols = n;
        q->cols = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        memcpy(q->cols, cols, sizeof(int) * (size_t)n);
        q->lq = lq; q->tau = tau; q->ntau = ntau; q->has = 1;
        res.b = b; res.u = u; res.D = D; res.U = U; res.V = V;
        res.cols = cols; res.ncols = n; res.nk = nk;
        return res;
    }
    const hsso_node* l = &c->nd[x->left];
    const hsso_node* r = &c->nd[x->right];
    int m1 = l->m, n1 = l->n;
    /* b1 = b[1:m1, :]; b2 = b[m1+1:end, :]  (:69) */
    mat b1 = msub(&b, 0, m1, 0, b.c);
    mat b2 = msub(&b, m1, b.r, 0, b.c);
    mfree(&b);
    ulvres r1 = ulv_rec(c, x->left, b1, co, 0);
    ulvres r2 = ulv_rec(c, x->right, b2, co + n1, 0);
    const double* B12 = c->blob + (x->off_B12 < 0 ? 0 : x->off_B12);
    const double* B21 = c->blob + (x->off_B21 < 0 ? 0 : x->off_B21);
    int nrhs = c->nrhs;
    /* U1*B12 and U2*B21 (left-to-right association, :74-75) */
    mat UB12 = mnew(r1.U.r, r->kv), UB21 = mnew(r2.U.r, l->kv);
    gemm(c->L, 'N', 'N', r1.U.r, r->kv, l->ku, 1.0, r1.U.a, r1.U.r, B12, l->ku, 0.0, UB12.a, r1.U.r);
    gemm(c->L, 'N', 'N', r2.U.r, l->kv, r->ku, 1.0, r2.U.a, r2.U.r, B21, r->ku, 0.0, UB21.a, r2.U.r);
    /* b = [b1; b2] .- [U1*B12*u2; U2*B21*u1]  (:74) */
    int mb1 = r1.b.r, mb2 = r2.b.r;
    mat bm = mnew(mb1 + mb2, nrhs);
    {
        mat t1 = mnew(mb1, nrhs), t2 = mnew(mb2, nrhs);
        gemm(c->L, 'N', 'N', mb1, nrhs, r->kv, 1.0, UB12.a, UB12.r, r2.u.a, r2.u.r, 0.0, t1.a, mb1);
        gemm(c->L, 'N', 'N', mb2, nrhs, l->kv, 1.0, UB21.a, UB21.r, r1.u.a, r1.u.r, 0.0, t2.a, mb2);
        for (int j = 0; j < nrhs; ++j) {
            for (int q = 0; q < mb1; ++q)
                bm.a[q + (size_t)j * (mb1 + mb2)] = r1.b.a[q + (size_t)j * mb1] - t1.a[q + (size_t)j * mb1];
            for (int q = 0; q < mb2; ++q)
                bm.a[mb1 + q + (size_t)j * (mb1 + mb2)] = r2.b.a[q + (size_t)j * mb2] - t2.a[q + (size_t)j * mb2];
        }
        mfree(&t1); mfree(&t2);
    }
    /* D = [D1 U1*B12*V2'; U2*B21*V1' D2]  (:75) */
    int md1 = r1.D.r, nd1 = r1.D.c, md2 = r2.D.r, nd2 = r2.D.c;
    mat D = mnew(md1 + md2, nd1 + nd2);
    {
        int M = md1 + md2;
        for (int j = 0; j < nd1; ++j)
            for (int q = 0; q < md1; ++q) D.a[q + (size_t)j * M] = r1.D.a[q + (size_t)j * md1];
        for (int j = 0; j < nd2; ++j)
            for (int q = 0; q < md2; ++q) D.a[md1 + q + (size_t)(nd1 + j) * M] = r2.D.a[q + (size_t)j * md2];
        gemm(c->L, 'N', 'T', md1, nd2, r->kv, 1.0, UB12.a, UB12.r, r2.V.a, r2.V.r, 0.0, D.a + (size_t)nd1 * M, M);
        gemm(c->L, 'N', 'T', md2, nd1, l->kv, 1.0, UB21.a, UB21.r, r1.V.a, r1.V.r, 0.0, D.a + md1, M);
    }
    /* cols = [cols1[nk1+1:end]; cols2[nk2+1:end]]  (:76) */
    int nc = (r1.ncols - r1.nk) + (r2.ncols - r2.nk);
    int* cols = (int*)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    memcpy(cols, r1.cols + r1.nk, sizeof(int) * (size_t)(r1.ncols - r1.nk));
    memcpy(cols + (r1.ncols - r1.nk), r2.cols + r2.nk, sizeof(int) * (size_t)(r2.ncols - r2.nk));
    /* U = [U1*R1; U2*R2]; V = [V1*W1; V2*W2]  (:78-79) */
    int ku = x->ku, kv = x->kv;
    mat U = mnew(r1.U.r + r2.U.r, ku), V = mnew(r1.V.r + r2.V.r, kv);
    gemm(c->L, 'N', 'N', r1.U.r, ku, l->ku, 1.0, r1.U.a, r1.U.r, c->blob + (l->off_R < 0 ? 0 : l->off_R), l->ku, 0.0,
         U.a, U.r);
    gemm(c->L, 'N', 'N', r2.U.r, ku, r->ku, 1.0, r2.U.a, r2.U.r, c->blob + (r->off_R < 0 ? 0 : r->off_R), r->ku, 0.0,
         U.a + r1.U.r, U.r);
    gemm(c->L, 'N', 'N', r1.V.r, kv, l->kv, 1.0, r1.V.a, r1.V.r, c->blob + (l->off_W < 0 ? 0 : l->off_W), l->kv, 0.0,
         V.a, V.r);
    gemm(c->L, 'N', 'N', r2.V.r, kv, r->kv, 1.0, r2.V.a, r2.V.r, c->blob + (r->off_W < 0 ? 0 : r->off_W), r->kv, 0.0,
         V.a + r1.V.r, V.r);
    mfree(&UB12); mfree(&UB21);

    if (rootnode) {
        /* z[cols, :] = D \ b  (:82-86) */
        int M = D.r;
        if (M != D.c || M != nc) { c->info = HSSO_DIM; }
        else if (M > 0) {
            int* ipiv = (int*)malloc(sizeof(int) * (size_t)M);
            int info = 0, lda = M;
            c->L->dgetrf(&M, &M, D.a, &lda, ipiv, &info);
            if (info > 0) c->info = HSSO_SINGULAR;
            else if (info < 0) c->info = HSSO_LAPACK;
            else {
                char tr = 'N';
                int nr = nrhs, ldb = M;
                if (nr > 0) c->L->dgetrs(&tr, &M, &nr, D.a, &lda, ipiv, bm.a, &ldb, &info);
                store_zloc(c, cols, M, &bm);
            }
            free(ipiv);
        }
        res.b = bm; res.u = mnew(0, nrhs); res.D = D; res.U = U; res.V = V;
        res.cols = cols; res.ncols = nc; res.nk = D.c;
        c->QV[i].has = 0;
    } else {
        mat zloc, u, lq; double* tau; int ntau, mk, nk;
        ulv_reduce(c, &D, &U, &V, &bm, &zloc, &u, &mk, &nk, &lq, &tau, &ntau);
        /* u .= u .+ W1'*u1 .+ W2'*u2  (:89) */
        {
            mat t1 = mnew(kv, nrhs), t2 = mnew(kv, nrhs);
            gemm(c->L, 'T', 'N', kv, nrhs, l->kv, 1.0, c->blob + (l->off_W < 0 ? 0 : l->off_W), l->kv, r1.u.a,
                 r1.u.r, 0.0, t1.a, kv);
            gemm(c->L, 'T', 'N', kv, nrhs, r->kv, 1.0, c->blob + (r->off_W < 0 ? 0 : r->off_W), r->kv, r2.u.a,
                 r2.u.r, 0.0, t2.a, kv);
            for (size_t e = 0; e < (size_t)kv * nrhs; ++e) u.a[e] = u.a[e] + t1.a[e] + t2.a[e];
            mfree(&t1); mfree(&t2);
        }
        store_zloc(c, cols, mk, &zloc);
        mfree(&zloc);
        qvnode* q = &c->QV[i];
        q->ncols = nc;
        q->cols = (int*)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
        memcpy(q->cols, cols, sizeof(int) * (size_t)nc);
        q->lq = lq; q->tau = tau; q->ntau = ntau; q->has = 1;
        res.b = bm; res.u = u; res.D = D; res.U = U; res.V = V;
        res.cols = cols; res.ncols = nc; res.nk = nk;
    }
    mfree(&r1.b); mfree(&r1.u); mfree(&r1.D); mfree(&r1.U); mfree(&r1.V); free(r1.cols);
    mfree(&r2.b); mfree(&r2.u); mfree(&r2.D); mfree(&r2.U); mfree(&r2.V); free(r2.cols);
    return res;
}

/* _ulvfactsolve_topdown!(QV, z), ulvfactor.jl:98-107 (pre-order) */
static void ulv_topdown(ulvctx* c, int i) {
    qvnode* q = &c->QV[i];
    const hsso_node* x = &c->nd[i];
    if (q->has && q->ntau > 0 && c->nrhs > 0) {
        int nc = q->ncols, nrhs = c->nrhs;
        mat zc = mnew(nc, nrhs);
        for (int j = 0; j < nrhs; ++j)
            for (int r = 0; r < nc; ++r) zc.a[r + (size_t)j * nc] = c->z[q->cols[r] + (size_t)j * c->ldz];
        if (orm(c->L->dormlq, 'L', 'T', nc, nrhs, q->ntau, q->lq.a, q->lq.r, q->tau, zc.a, nc)) c->info = HSSO_LAPACK;
        for (int j = 0; j < nrhs; ++j)
            for (int r = 0; r < nc; ++r) c->z[q->cols[r] + (size_t)j * c->ldz] = zc.a[r + (size_t)j * nc];
        mfree(&zc);
    }
    if (x->left >= 0) ulv_topdown(c, x->left);
    if (x->right >= 0) ulv_topdown(c, x->right);
}

/* ulvfactsolve(hssA, b), ulvfactor.jl:10-19: z = A \ b for square HSS A.
 * b is n x nrhs (ld ldb), z is n x nrhs (ld ldz).  Returns HSSO_* status. */
int hsso_ulvfactsolve(const hsso_lapack* L, const hsso_node* nodes, int nn, const double* blob, int nrhs,
                      const double* b, int ldb, double* z, int ldz) {
    if (nn <= 0 || nrhs < 0) return HSSO_DIM;
    const hsso_node* root = &nodes[0];
    int n = root->n;
    if (root->m != root->n) return HSSO_DIM;
    if (nrhs == 0) return HSSO_OK;
    if (root->left < 0) {
        /* hssA.D \ b  (:11-12) */
        mat D = mview_blob(blob, root->off_D, n, n);
        int* ipiv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        int info = 0, lda = n > 0 ? n : 1;
        L->dgetrf(&n, &n, D.a, &lda, ipiv, &info);
        int st = HSSO_OK;
        if (info > 0) st = HSSO_SINGULAR;
        else if (info < 0) st = HSSO_LAPACK;
        else {
            for (int j = 0; j < nrhs; ++j) memcpy(z + (size_t)j * ldz, b + (size_t)j * ldb, sizeof(double) * (size_t)n);
            char tr = 'N';
            int nr = nrhs, ld = ldz;
            L->dgetrs(&tr, &n, &nr, D.a, &lda, ipiv, z, &ld, &info);
        }
        free(ipiv);
        mfree(&D);
        return st;
    }
    ulvctx c;
    memset(&c, 0, sizeof(c));
    c.L = L; c.nd = nodes; c.blob = blob; c.nrhs = nrhs; c.z = z; c.ldz = ldz;
    c.QV = (qvnode*)calloc((size_t)nn, sizeof(qvnode));
    if (!c.QV) return HSSO_OOM;
    for (int j = 0; j < nrhs; ++j) memset(z + (size_t)j * ldz, 0, sizeof(double) * (size_t)n); /* :14 */
    mat bb = mcopy(b, n, nrhs, ldb);
    ulvres r = ulv_rec(&c, 0, bb, 0, 1);
    mfree(&r.b); mfree(&r.u); mfree(&r.D); mfree(&r.U); mfree(&r.V); free(r.cols);
    if (c.info == HSSO_OK) ulv_topdown(&c, 0);
    for (int i = 0; i < nn; ++i) {
        free(c.QV[i].cols);
        mfree(&c.QV[i].lq);
        free(c.QV[i].tau);
    }
    free(c.QV);
    return c.info;
}
