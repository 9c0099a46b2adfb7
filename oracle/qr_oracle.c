/*
 * qr_oracle.c -- CPU restatement of qrusty's rowwise Pauli-operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library under qrusty_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  Nothing in the product links it.
 *
 * Each function restates one piece of the reference algorithm and cites the
 * reference file:line it follows (paths relative to the qrusty source tree).
 * Build with -ffp-contract=off: the reference is Rust, which never contracts
 * a*b+c into an FMA, and the complex products below must round the same way.
 *
 * Parity status of each function (how it is pinned; see tests/test_oracle.py):
 *   orc_parse_label        KATs of qrusty/src/lib.rs:608-693 + Pauli matrices
 *   orc_make_params        closed form (-i)^phase * c, checked on H2/H4/H6 by a
 *                          dense Kronecker build (numpy.kron, a library routine)
 *   orc_make_row / build   dense Kronecker equality on H2/H4/H6 (the assertion
 *                          of pyqrusty/tests/test_H.py:32-48), closed-form
 *                          structure (indptr = r*G, sorted r^x_g), brute force
 *   orc_spmat_dot_densevec scipy csr_matrix.dot and dense matmul (tolerance);
 *                          bit order inside sprs dot_dense is third-party code
 *                          that is absent here: "parity unpinned" at bit level
 *   orc_apply_rows         equals CSR build + matvec (checked on small cases)
 *   orc_csr_diagonal       numpy.diagonal of the dense matrix; test_it.py:46-51
 *   orc_axpby              bit-exact against CPython complex * and + (IEEE
 *                          double, no contraction), allclose as test_it.py:232-263
 *   orc_precond2           CPython complex division (Smith's method) within a
 *                          few ulp; reg threshold cases; the H2 vector printed
 *                          in test_it.py:285-300 (to its 9 digits).  The bit
 *                          order of num-complex's division is not available
 *                          here: "parity unpinned" at the last bit
 *   orc_count_zeros,       numpy's abs (hypot) on random data, the
 *   orc_eliminate_zeros    reference's test_nz (test_it.py:140-147) and
 *                          scipy's CSR of the kept entries
 *   orc_csr_add            scipy's sparse a*A + b*B (which also drops exact
 *                          zeros) on random and fixture matrices, and the
 *                          dense sum.  Whether sprs's csmat_binop drops exact
 *                          zeros is not available here (the crate is absent):
 *                          "parity unpinned" for that one rule
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <stdatomic.h>

typedef struct { double re, im; } orc_c128;

/* ---------------------------------------------------------------------------
 * Label parsing: qrusty/src/lib.rs:125-149 (regex ^([+-]?)1?([ij]?)([IXYZ]+)$,
 * characters reversed so qubit q is label[len-1-q]), x/z bits lib.rs:80-91,
 * masks lib.rs:161-176, phase lib.rs:179-181.
 * Returns 0 on success, -1 on a malformed label, -2 if more than 63 qubits.
 * ------------------------------------------------------------------------- */
int orc_parse_label(const char *s, uint32_t *nq, uint32_t *base_phase,
                    uint64_t *x, uint64_t *z, uint32_t *phase)
{
    const char *p = s;
    int neg = 0, imag = 0;
    if (*p == '+' || *p == '-') { neg = (*p == '-'); p++; }
    if (*p == '1') p++;
    if (*p == 'i' || *p == 'j') { imag = 1; p++; }
    size_t n = strlen(p);
    if (n == 0) return -1;
    for (size_t k = 0; k < n; k++)
        if (p[k] != 'I' && p[k] != 'X' && p[k] != 'Y' && p[k] != 'Z') return -1;
    if (n > 63) return -2;
    uint32_t bp = (uint32_t)(imag + (neg ? 2 : 0));      /* lib.rs:136-141 */
    uint64_t xm = 0, zm = 0;
    uint32_t ny = 0;
    for (size_t q = 0; q < n; q++) {
        char c = p[n - 1 - q];                           /* lib.rs:144 */
        if (c == 'X' || c == 'Y') xm |= (uint64_t)1 << q;
        if (c == 'Y' || c == 'Z') zm |= (uint64_t)1 << q;
        if (c == 'Y') ny++;
    }
    *nq = (uint32_t)n;
    *base_phase = bp;
    *x = xm;
    *z = zm;
    *phase = (bp + ny) % 4;                              /* lib.rs:179-181 */
    return 0;
}

/* num-complex 0.4 Complex64 multiply: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline orc_c128 cmul(orc_c128 a, orc_c128 b)
{
    orc_c128 r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* ---------------------------------------------------------------------------
 * make_params: qrusty/src/accel.rs:141-157.  c' = M[phase] * c with
 * M = {1, -i, -1, +i}; the multiply is the num-complex product above.
 * phase_in[t] is (base_phase + #Y) mod 4 as returned by orc_parse_label.
 * ------------------------------------------------------------------------- */
void orc_make_params(uint64_t T, const uint32_t *phase_in, const double *coeff_in,
                     double *coeff_out)
{
    static const orc_c128 M[4] = {{1.0, 0.0}, {0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}};
    for (uint64_t t = 0; t < T; t++) {
        orc_c128 c = {coeff_in[2 * t], coeff_in[2 * t + 1]};
        orc_c128 r = cmul(M[phase_in[t] % 4], c);
        coeff_out[2 * t] = r.re;
        coeff_out[2 * t + 1] = r.im;
    }
}

/* ---------------------------------------------------------------------------
 * make_row: qrusty/src/accel.rs:171-210.
 *   for each term: col = row ^ x_t, val = popcount(row & z_t) odd ? -c' : c'
 *   stable sort by col (Rust sort_by is a stable merge sort, accel.rs:188)
 *   fold equal cols: first value, then running + next (accel.rs:191-208);
 *   exact zeros are kept.
 * A stable sort's output is unique, so any stable sort restates sort_by.
 * Writes at most T entries to cols/vals and returns the count.
 * tmp must hold 2*T entries of (col, val).
 * ------------------------------------------------------------------------- */
typedef struct { uint64_t col; orc_c128 val; } orc_entry;

static void merge_sort_stable(orc_entry *a, orc_entry *buf, size_t n)
{
    /* bottom-up merge sort; ties keep the left run first => stable */
    orc_entry *src = a, *dst = buf;
    for (size_t width = 1; width < n; width *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * width) {
            size_t mid = lo + width < n ? lo + width : n;
            size_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            size_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (src[j].col < src[i].col) dst[k++] = src[j++];
                else dst[k++] = src[i++];
            }
            while (i < mid) dst[k++] = src[i++];
            while (j < hi) dst[k++] = src[j++];
        }
        orc_entry *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(orc_entry));
}

static size_t make_row_impl(uint64_t T, const uint64_t *xs, const uint64_t *zs,
                            const double *cp, uint64_t row, orc_entry *tmp,
                            uint64_t *cols, double *vals)
{
    orc_entry *v = tmp, *buf = tmp + T;
    for (uint64_t t = 0; t < T; t++) {
        orc_c128 c = {cp[2 * t], cp[2 * t + 1]};
        orc_c128 m = {-c.re, -c.im};                      /* accel.rs:178 */
        v[t].col = row ^ xs[t];
        v[t].val = (__builtin_popcountll(row & zs[t]) % 2 == 1) ? m : c;
    }
    merge_sort_stable(v, buf, (size_t)T);
    size_t out = 0;
    uint64_t col = v[0].col;
    orc_c128 sum = v[0].val;
    for (uint64_t t = 1; t < T; t++) {
        if (v[t].col == col) {
            sum.re = sum.re + v[t].val.re;
            sum.im = sum.im + v[t].val.im;
        } else {
            cols[out] = col; vals[2 * out] = sum.re; vals[2 * out + 1] = sum.im; out++;
            col = v[t].col;
            sum = v[t].val;
        }
    }
    cols[out] = col; vals[2 * out] = sum.re; vals[2 * out + 1] = sum.im; out++;
    return out;
}

size_t orc_make_row(uint64_t T, const uint64_t *xs, const uint64_t *zs,
                    const double *cp, uint64_t row, uint64_t *cols, double *vals)
{
    orc_entry *tmp = (orc_entry *)malloc(2 * T * sizeof(orc_entry));
    if (!tmp) return 0;
    size_t n = make_row_impl(T, xs, zs, cp, row, tmp, cols, vals);
    free(tmp);
    return n;
}

/* ---------------------------------------------------------------------------
 * A tiny dynamic-scheduling thread pool (rayon's par_iter over chunks).
 * ------------------------------------------------------------------------- */
typedef struct {
    atomic_ullong next;
    uint64_t nchunks;
    void (*fn)(void *ctx, uint64_t chunk);
    void *ctx;
} orc_pool;

static void *pool_worker(void *arg)
{
    orc_pool *p = (orc_pool *)arg;
    for (;;) {
        uint64_t c = atomic_fetch_add(&p->next, 1);
        if (c >= p->nchunks) break;
        p->fn(p->ctx, c);
    }
    return NULL;
}

static void pool_run(int nthreads, uint64_t nchunks, void (*fn)(void *, uint64_t), void *ctx)
{
    orc_pool p;
    atomic_init(&p.next, 0);
    p.nchunks = nchunks;
    p.fn = fn;
    p.ctx = ctx;
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > nchunks) nthreads = (int)(nchunks ? nchunks : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, pool_worker, &p);
    pool_worker(&p);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
}

/* ---------------------------------------------------------------------------
 * make_unsafe_vectors_chunked: qrusty/src/accel.rs:267-336, restated with its
 * memory pattern: each row allocates its (col, val) vector and its two output
 * vectors (accel.rs:174-190), each chunk of `step` rows concatenates its rows
 * into chunk buffers (accel.rs:293-301), a serial scan builds indptr
 * (accel.rs:309-319) and the chunk buffers are concatenated into the final
 * arrays in parallel (accel.rs:322-325).
 * Rows [row_begin, row_end) are built; indptr has (row_end-row_begin+1)
 * entries counted from 0.  Returns nnz, or UINT64_MAX if capacity would be
 * exceeded (cap = capacity of indices/data in entries).
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t T; const uint64_t *xs, *zs; const double *cp;
    uint64_t row_begin, row_end, step;
    uint64_t **chunk_nnz;      /* per chunk: per-row counts */
    uint64_t **chunk_cols;
    double **chunk_vals;
    uint64_t *chunk_total;
} build_ctx;

static void build_chunk(void *vctx, uint64_t c)
{
    build_ctx *b = (build_ctx *)vctx;
    uint64_t lo = b->row_begin + c * b->step;
    uint64_t hi = lo + b->step < b->row_end ? lo + b->step : b->row_end;
    uint64_t nrows = hi - lo;
    uint64_t **rcols = (uint64_t **)malloc(nrows * sizeof(uint64_t *));
    double **rvals = (double **)malloc(nrows * sizeof(double *));
    uint64_t *rn = (uint64_t *)malloc(nrows * sizeof(uint64_t));
    uint64_t total = 0;
    for (uint64_t r = lo; r < hi; r++) {
        orc_entry *tmp = (orc_entry *)malloc(2 * b->T * sizeof(orc_entry));
        uint64_t *cols = (uint64_t *)malloc(b->T * sizeof(uint64_t));
        double *vals = (double *)malloc(2 * b->T * sizeof(double));
        size_t n = make_row_impl(b->T, b->xs, b->zs, b->cp, r, tmp, cols, vals);
        free(tmp);
        rcols[r - lo] = cols; rvals[r - lo] = vals; rn[r - lo] = n;
        total += n;
    }
    uint64_t *cc = (uint64_t *)malloc((total ? total : 1) * sizeof(uint64_t));
    double *cv = (double *)malloc((total ? total : 1) * 2 * sizeof(double));
    uint64_t off = 0;
    for (uint64_t i = 0; i < nrows; i++) {
        memcpy(cc + off, rcols[i], rn[i] * sizeof(uint64_t));
        memcpy(cv + 2 * off, rvals[i], rn[i] * 2 * sizeof(double));
        off += rn[i];
        free(rcols[i]); free(rvals[i]);
    }
    free(rcols); free(rvals);
    b->chunk_nnz[c] = rn;
    b->chunk_cols[c] = cc;
    b->chunk_vals[c] = cv;
    b->chunk_total[c] = total;
}

typedef struct {
    build_ctx *b; uint64_t *offs; uint64_t *indices; double *data;
} concat_ctx;

static void concat_chunk(void *vctx, uint64_t c)
{
    concat_ctx *k = (concat_ctx *)vctx;
    uint64_t n = k->b->chunk_total[c];
    memcpy(k->indices + k->offs[c], k->b->chunk_cols[c], n * sizeof(uint64_t));
    memcpy(k->data + 2 * k->offs[c], k->b->chunk_vals[c], n * 2 * sizeof(double));
    free(k->b->chunk_cols[c]);
    free(k->b->chunk_vals[c]);
}

uint64_t orc_build_rows_chunked(uint64_t T, const uint64_t *xs, const uint64_t *zs,
                                const double *cp, uint64_t row_begin, uint64_t row_end,
                                uint64_t step, int nthreads, uint64_t cap,
                                uint64_t *indptr, uint64_t *indices, double *data)
{
    if (step == 0 || row_end < row_begin || T == 0) return UINT64_MAX;
    uint64_t nrows = row_end - row_begin;
    uint64_t nchunks = (nrows + step - 1) / step;
    build_ctx b;
    b.T = T; b.xs = xs; b.zs = zs; b.cp = cp;
    b.row_begin = row_begin; b.row_end = row_end; b.step = step;
    b.chunk_nnz = (uint64_t **)calloc(nchunks ? nchunks : 1, sizeof(uint64_t *));
    b.chunk_cols = (uint64_t **)calloc(nchunks ? nchunks : 1, sizeof(uint64_t *));
    b.chunk_vals = (double **)calloc(nchunks ? nchunks : 1, sizeof(double *));
    b.chunk_total = (uint64_t *)calloc(nchunks ? nchunks : 1, sizeof(uint64_t));
    pool_run(nthreads, nchunks, build_chunk, &b);

    /* serial indptr scan (accel.rs:309-319) */
    uint64_t nnz = 0;
    for (uint64_t r = 0; r < nrows; r++) {
        uint64_t chunkind = r / step, chunkofs = r % step;
        indptr[r] = nnz;
        nnz += b.chunk_nnz[chunkind][chunkofs];
    }
    indptr[nrows] = nnz;
    uint64_t ret = nnz;
    uint64_t *offs = (uint64_t *)malloc((nchunks ? nchunks : 1) * sizeof(uint64_t));
    uint64_t acc = 0;
    for (uint64_t c = 0; c < nchunks; c++) { offs[c] = acc; acc += b.chunk_total[c]; }
    if (nnz > cap) {
        for (uint64_t c = 0; c < nchunks; c++) { free(b.chunk_cols[c]); free(b.chunk_vals[c]); }
        ret = UINT64_MAX;
    } else {
        concat_ctx k = {&b, offs, indices, data};
        pool_run(nthreads, nchunks, concat_chunk, &k);   /* accel.rs:322-325 */
    }
    for (uint64_t c = 0; c < nchunks; c++) free(b.chunk_nnz[c]);
    free(offs);
    free(b.chunk_nnz); free(b.chunk_cols); free(b.chunk_vals); free(b.chunk_total);
    return ret;
}

/* ---------------------------------------------------------------------------
 * spmat_dot_densevec: qrusty/src/accel.rs:338-370.  Chunks of 1024 rows run in
 * parallel; each row is sprs CsVecView::dot_dense (accel.rs:359), restated as
 * a fold from zero of data[k]*x[indices[k]] in k order (num-complex Sum) with
 * the num-complex product; chunk outputs are then concatenated.
 * The inner order lives in sprs (third-party, absent): parity unpinned at the
 * bit level, pinned to tolerance by scipy and dense products in the tests.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t nrows; const uint64_t *indptr, *indices; const double *data, *x;
    double **chunk_out;
} spmv_ctx;

static void spmv_chunk(void *vctx, uint64_t c)
{
    spmv_ctx *s = (spmv_ctx *)vctx;
    uint64_t lo = c * 1024, hi = lo + 1024 < s->nrows ? lo + 1024 : s->nrows;
    double *w = (double *)malloc((hi - lo) * 2 * sizeof(double));
    for (uint64_t r = lo; r < hi; r++) {
        orc_c128 acc = {0.0, 0.0};
        for (uint64_t k = s->indptr[r]; k < s->indptr[r + 1]; k++) {
            orc_c128 a = {s->data[2 * k], s->data[2 * k + 1]};
            uint64_t j = s->indices[k];
            orc_c128 b = {s->x[2 * j], s->x[2 * j + 1]};
            orc_c128 p = cmul(a, b);
            acc.re = acc.re + p.re;
            acc.im = acc.im + p.im;
        }
        w[2 * (r - lo)] = acc.re;
        w[2 * (r - lo) + 1] = acc.im;
    }
    s->chunk_out[c] = w;
}

void orc_spmat_dot_densevec(uint64_t nrows, const uint64_t *indptr, const uint64_t *indices,
                            const double *data, const double *x, double *y, int nthreads)
{
    uint64_t nchunks = (nrows + 1023) / 1024;
    spmv_ctx s = {nrows, indptr, indices, data, x,
                  (double **)calloc(nchunks ? nchunks : 1, sizeof(double *))};
    pool_run(nthreads, nchunks, spmv_chunk, &s);
    for (uint64_t c = 0; c < nchunks; c++) {          /* unsafe_par_concat_slices */
        uint64_t lo = c * 1024, hi = lo + 1024 < nrows ? lo + 1024 : nrows;
        memcpy(y + 2 * lo, s.chunk_out[c], (hi - lo) * 2 * sizeof(double));
        free(s.chunk_out[c]);
    }
    free(s.chunk_out);
}

/* ---------------------------------------------------------------------------
 * Matrix-free apply on selected rows, by construction equal to the CSR build
 * followed by spmat_dot_densevec: for each requested row r it runs make_row
 * (accel.rs:171-210) and then the dot of accel.rs:359 over the row entries.
 * nvec state vectors are stored one after another (N entries each).
 * out[(v*nrows_sel + i)] = (A x_v)[rows[i]].
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t T; const uint64_t *xs, *zs; const double *cp;
    const uint64_t *rows; uint64_t nsel; uint64_t N; uint32_t nvec;
    const double *x; double *out;
} apply_ctx;

#define APPLY_CHUNK 256
static void apply_chunk(void *vctx, uint64_t c)
{
    apply_ctx *a = (apply_ctx *)vctx;
    orc_entry *tmp = (orc_entry *)malloc(2 * a->T * sizeof(orc_entry));
    uint64_t *cols = (uint64_t *)malloc(a->T * sizeof(uint64_t));
    double *vals = (double *)malloc(2 * a->T * sizeof(double));
    uint64_t lo = c * APPLY_CHUNK, hi = lo + APPLY_CHUNK < a->nsel ? lo + APPLY_CHUNK : a->nsel;
    for (uint64_t i = lo; i < hi; i++) {
        uint64_t r = a->rows[i];
        size_t n = make_row_impl(a->T, a->xs, a->zs, a->cp, r, tmp, cols, vals);
        for (uint32_t v = 0; v < a->nvec; v++) {
            const double *xv = a->x + 2 * (uint64_t)v * a->N;
            orc_c128 acc = {0.0, 0.0};
            for (size_t k = 0; k < n; k++) {
                orc_c128 d = {vals[2 * k], vals[2 * k + 1]};
                orc_c128 b = {xv[2 * cols[k]], xv[2 * cols[k] + 1]};
                orc_c128 p = cmul(d, b);
                acc.re = acc.re + p.re;
                acc.im = acc.im + p.im;
            }
            a->out[2 * ((uint64_t)v * a->nsel + i)] = acc.re;
            a->out[2 * ((uint64_t)v * a->nsel + i) + 1] = acc.im;
        }
    }
    free(tmp); free(cols); free(vals);
}

void orc_apply_rows(uint64_t T, const uint64_t *xs, const uint64_t *zs, const double *cp,
                    uint64_t N, uint32_t nvec, const double *x,
                    const uint64_t *rows, uint64_t nsel, double *out, int nthreads)
{
    apply_ctx a = {T, xs, zs, cp, rows, nsel, N, nvec, x, out};
    pool_run(nthreads, (nsel + APPLY_CHUNK - 1) / APPLY_CHUNK, apply_chunk, &a);
}

/* ---------------------------------------------------------------------------
 * Full-state FP64 apply through grouped terms, the reference semantics
 * reorganised for speed (no per-row sort): the terms of each distinct x mask
 * are folded in term order (first value, then +), exactly as make_row folds
 * them after its stable sort (accel.rs:188-208), so D_g(r) is bit-identical
 * to the CSR value of group g in row r.  The row's groups are then accumulated
 * in ascending column order from zero (accel.rs:359), which needs the column
 * order r ^ x_g sorted: done here with a per-row sort of G keys.
 * Used for full-size FP64 references (e.g. <psi|H|psi> of a 2^24 state).
 * gx[G] distinct masks, gstart[G+1] ranges into term arrays grouped by mask
 * (terms of one group in original order).  Returns y for all rows.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t G; const uint64_t *gx; const uint64_t *gstart;
    const uint64_t *tz; const double *tc;
    uint64_t N; uint32_t nvec; const double *x; double *y;
} gapply_ctx;

typedef struct { uint64_t col; uint64_t g; } colg;

static int colg_cmp(const void *a, const void *b)
{
    uint64_t x = ((const colg *)a)->col, y = ((const colg *)b)->col;
    return (x > y) - (x < y);
}

#define GAPPLY_CHUNK 1024
static void gapply_chunk(void *vctx, uint64_t c)
{
    gapply_ctx *a = (gapply_ctx *)vctx;
    colg *order = (colg *)malloc(a->G * sizeof(colg));
    orc_c128 *D = (orc_c128 *)malloc(a->G * sizeof(orc_c128));
    uint64_t lo = c * GAPPLY_CHUNK, hi = lo + GAPPLY_CHUNK < a->N ? lo + GAPPLY_CHUNK : a->N;
    for (uint64_t r = lo; r < hi; r++) {
        for (uint64_t g = 0; g < a->G; g++) {
            orc_c128 s = {0, 0};
            for (uint64_t t = a->gstart[g]; t < a->gstart[g + 1]; t++) {
                orc_c128 cv = {a->tc[2 * t], a->tc[2 * t + 1]};
                if (__builtin_popcountll(r & a->tz[t]) & 1) { cv.re = -cv.re; cv.im = -cv.im; }
                if (t == a->gstart[g]) s = cv;
                else { s.re = s.re + cv.re; s.im = s.im + cv.im; }
            }
            D[g] = s;
            order[g].col = r ^ a->gx[g];
            order[g].g = g;
        }
        qsort(order, a->G, sizeof(colg), colg_cmp);
        for (uint32_t v = 0; v < a->nvec; v++) {
            const double *xv = a->x + 2 * (uint64_t)v * a->N;
            orc_c128 acc = {0.0, 0.0};
            for (uint64_t k = 0; k < a->G; k++) {
                orc_c128 b = {xv[2 * order[k].col], xv[2 * order[k].col + 1]};
                orc_c128 p = cmul(D[order[k].g], b);
                acc.re = acc.re + p.re;
                acc.im = acc.im + p.im;
            }
            a->y[2 * ((uint64_t)v * a->N + r)] = acc.re;
            a->y[2 * ((uint64_t)v * a->N + r) + 1] = acc.im;
        }
    }
    free(order); free(D);
}

void orc_apply_grouped(uint64_t G, const uint64_t *gx, const uint64_t *gstart,
                       const uint64_t *tz, const double *tc, uint64_t N, uint32_t nvec,
                       const double *x, double *y, int nthreads)
{
    gapply_ctx a = {G, gx, gstart, tz, tc, N, nvec, x, y};
    pool_run(nthreads, (N + GAPPLY_CHUNK - 1) / GAPPLY_CHUNK, gapply_chunk, &a);
}

/* <psi|y> = sum_r conj(psi_r) * y_r in FP64, in row order, per vector. */
void orc_vdot(uint64_t N, uint32_t nvec, const double *psi, const double *y, double *out)
{
    for (uint32_t v = 0; v < nvec; v++) {
        double re = 0.0, im = 0.0;
        const double *p = psi + 2 * (uint64_t)v * N, *q = y + 2 * (uint64_t)v * N;
        for (uint64_t r = 0; r < N; r++) {
            re = re + (p[2 * r] * q[2 * r] + p[2 * r + 1] * q[2 * r + 1]);
            im = im + (p[2 * r] * q[2 * r + 1] - p[2 * r + 1] * q[2 * r]);
        }
        out[2 * v] = re;
        out[2 * v + 1] = im;
    }
}

/* ---------------------------------------------------------------------------
 * Eigensolver-step operations (pyqrusty/src/lib.rs:118-125, 436-468;
 * qrusty/src/accel.rs:374-393).  num-complex 0.4 arithmetic: Complex * is
 * (a.re x.re - a.im x.im, a.re x.im + a.im x.re); Complex / is
 * ((a c + b d) / (c c + d d), (b c - a d) / (c c + d d)); norm() is hypot.
 * ------------------------------------------------------------------------- */

/* sprs CsMat::diag().to_dense(): entry (r, r) of each row, 0 where absent;
 * rows hold ascending columns (a valid CsMat), searched in order here. */
void orc_csr_diagonal(uint64_t nrows, uint64_t ncols, const uint64_t *indptr,
                      const uint64_t *indices, const double *data, double *out)
{
    uint64_t n = nrows < ncols ? nrows : ncols;
    for (uint64_t r = 0; r < n; ++r) {
        out[2 * r] = 0.0;
        out[2 * r + 1] = 0.0;
        for (uint64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
            if (indices[k] == r) {
                out[2 * r] = data[2 * k];
                out[2 * r + 1] = data[2 * k + 1];
                break;
            }
        }
    }
}

/* accel.rs:374-393: mode 0 z = a x + b y, mode 1 z = a x + y, mode 2 z = a x */
void orc_axpby(uint64_t n, int mode, const double *a, const double *x, const double *b,
               const double *y, double *z)
{
    for (uint64_t i = 0; i < n; ++i) {
        double axr = a[0] * x[2 * i] - a[1] * x[2 * i + 1];
        double axi = a[0] * x[2 * i + 1] + a[1] * x[2 * i];
        if (mode == 2) {
            z[2 * i] = axr;
            z[2 * i + 1] = axi;
            continue;
        }
        double byr, byi;
        if (mode == 0) {
            byr = b[0] * y[2 * i] - b[1] * y[2 * i + 1];
            byi = b[0] * y[2 * i + 1] + b[1] * y[2 * i];
        } else {
            byr = y[2 * i];
            byi = y[2 * i + 1];
        }
        z[2 * i] = axr + byr;
        z[2 * i + 1] = axi + byi;
    }
}

/* pyqrusty/src/lib.rs:436-468: out = dx / reg(diag - e, tol), reg(v) = tol + 0i
 * when |v| < tol, else v */
void orc_precond2(uint64_t n, const double *diag, const double *dx, const double *e, double tol,
                  double *out)
{
    for (uint64_t i = 0; i < n; ++i) {
        double vr = diag[2 * i] - e[0], vi = diag[2 * i + 1] - e[1];
        if (hypot(vr, vi) < tol) {
            vr = tol;
            vi = 0.0;
        }
        double nn = vr * vr + vi * vi;
        double re = dx[2 * i] * vr + dx[2 * i + 1] * vi;
        double im = dx[2 * i + 1] * vr - dx[2 * i] * vi;
        out[2 * i] = re / nn;
        out[2 * i + 1] = im / nn;
    }
}

/* a*A + b*B (pyqrusty/src/lib.rs:499-517; sprs scale and csmat_binop): per row
 * the merge of the ascending columns; both: a*v1 + b*v2, one side only:
 * a*v1 + 0 or 0 + b*v2; exact zeros dropped.  With out_indices == NULL only
 * counts: out_indptr[r+1] - out_indptr[r] is filled and the total returned. */
uint64_t orc_csr_add(uint64_t nrows, const double *a, const uint64_t *ip1, const uint64_t *ix1,
                     const double *d1, const double *b, const uint64_t *ip2, const uint64_t *ix2,
                     const double *d2, uint64_t *out_indptr, uint64_t *out_indices,
                     double *out_data)
{
    uint64_t nnz = 0;
    out_indptr[0] = 0;
    for (uint64_t r = 0; r < nrows; ++r) {
        uint64_t i = ip1[r], i1 = ip1[r + 1], j = ip2[r], j1 = ip2[r + 1];
        while (i < i1 || j < j1) {
            uint64_t col;
            double vr, vi;
            if (j >= j1 || (i < i1 && ix1[i] < ix2[j])) {
                col = ix1[i];
                double pr = a[0] * d1[2 * i] - a[1] * d1[2 * i + 1];
                double pi = a[0] * d1[2 * i + 1] + a[1] * d1[2 * i];
                vr = pr + 0.0;
                vi = pi + 0.0;
                ++i;
            } else if (i >= i1 || ix2[j] < ix1[i]) {
                col = ix2[j];
                double qr = b[0] * d2[2 * j] - b[1] * d2[2 * j + 1];
                double qi = b[0] * d2[2 * j + 1] + b[1] * d2[2 * j];
                vr = 0.0 + qr;
                vi = 0.0 + qi;
                ++j;
            } else {
                col = ix1[i];
                double pr = a[0] * d1[2 * i] - a[1] * d1[2 * i + 1];
                double pi = a[0] * d1[2 * i + 1] + a[1] * d1[2 * i];
                double qr = b[0] * d2[2 * j] - b[1] * d2[2 * j + 1];
                double qi = b[0] * d2[2 * j + 1] + b[1] * d2[2 * j];
                vr = pr + qr;
                vi = pi + qi;
                ++i;
                ++j;
            }
            if (vr != 0.0 || vi != 0.0) {
                if (out_indices) {
                    out_indices[nnz] = col;
                    out_data[2 * nnz] = vr;
                    out_data[2 * nnz + 1] = vi;
                }
                ++nnz;
            }
        }
        out_indptr[r + 1] = nnz;
    }
    return nnz;
}

/* count_zeros / eliminate_zeros (qrusty/src/util.rs:144-171): |v| is
 * num-complex's norm(), hypot(re, im).  eliminate writes the kept entries in
 * order (rows of a CSR with ascending columns stay as they are) and returns
 * their number; out_indptr has nrows + 1 entries. */
uint64_t orc_count_zeros(uint64_t nnz, const double *d, double tol)
{
    uint64_t n = 0;
    for (uint64_t i = 0; i < nnz; ++i) n += hypot(d[2 * i], d[2 * i + 1]) <= tol;
    return n;
}

uint64_t orc_eliminate_zeros(uint64_t nrows, const uint64_t *ip, const uint64_t *ix, const double *d,
                             double tol, uint64_t *out_indptr, uint64_t *out_indices,
                             double *out_data)
{
    uint64_t k = 0;
    out_indptr[0] = 0;
    for (uint64_t r = 0; r < nrows; ++r) {
        for (uint64_t i = ip[r]; i < ip[r + 1]; ++i) {
            if (hypot(d[2 * i], d[2 * i + 1]) > tol) {
                out_indices[k] = ix[i];
                out_data[2 * k] = d[2 * i];
                out_data[2 * k + 1] = d[2 * i + 1];
                ++k;
            }
        }
        out_indptr[r + 1] = k;
    }
    return k;
}
