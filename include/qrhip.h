/*
 * qrhip.h -- C ABI of the MI355X (gfx950) Pauli-operator library libqrhip.so.
 *
 * The library replaces qrusty's rowwise CSR build and CSR matvec
 * (qrusty/src/accel.rs:141-370, reached from qrusty/src/lib.rs:552-585 and
 * pyqrusty/src/lib.rs:406-415, 476-491) with hand-written HIP kernels, and
 * adds a matrix-free Pauli-sum apply, <psi|H|psi>, and a batched launch.
 *
 * Conventions
 *  - Every function returns 0 (QR_OK) or a negative QR_ERR_* code; the message
 *    of the last failure on the calling thread is returned by qr_last_error().
 *    No C++ exception crosses this boundary.
 *  - Device pointers are plain device addresses on the current HIP device.
 *    Complex values are interleaved (re, im): complex128 = 2 doubles,
 *    complex64 = 2 floats.  Indices and indptr are uint64, as in the
 *    reference's CsMatI<Complex64, u64, u64> (qrusty/src/lib.rs:562-570).
 *  - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *    stream-ordered and asynchronous unless stated otherwise.
 *  - A qr_op may be used from several threads for launches; creating and
 *    destroying it is not concurrent with its use.
 */
#ifndef QRHIP_H
#define QRHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QRHIP_ABI_VERSION 1

enum qr_status {
    QR_OK = 0,
    QR_ERR_ARG = -1,      /* invalid argument (null pointer, bad size, ...)      */
    QR_ERR_LABEL = -2,    /* malformed Pauli label (lib.rs:129)                 */
    QR_ERR_TERMS = -3,    /* SparsePauliOp::new validation (lib.rs:354-376)     */
    QR_ERR_MODE = -4,     /* unrecognized AccelMode string (lib.rs:327-329)     */
    QR_ERR_SIZE = -5,     /* problem too large for the requested operation      */
    QR_ERR_HIP = -6,      /* a HIP runtime call failed                          */
    QR_ERR_NOMEM = -7,    /* allocation failed                                  */
    QR_ERR_IO = -8,       /* a file could not be opened, read or written        */
    QR_ERR_FORMAT = -9    /* a file's content is malformed                      */
};

/* Message of the last failure on this thread ("" if none). */
const char *qr_last_error(void);

/* QRHIP_ABI_VERSION of the loaded library. */
int qr_abi_version(void);

/* ------------------------------------------------------------------------
 * Device memory for callers without a HIP binding of their own (the qrusty
 * side holds Vec buffers; these give it the device counterparts, the role of
 * the qr_alloc / qr_free pair of the boundary).  Buffers belong to the current
 * device.  qr_copy_to_host synchronises the stream; qr_copy_to_device is
 * stream-ordered and synchronises too when `src` is pageable host memory.
 * ---------------------------------------------------------------------- */
int qr_device_alloc(uint64_t bytes, void **d_out);
int qr_device_free(void *d_ptr);
int qr_copy_to_device(void *d_dst, const void *src, uint64_t bytes, void *stream);
int qr_copy_to_host(void *dst, const void *d_src, uint64_t bytes, void *stream);

/* ------------------------------------------------------------------------
 * Pauli labels.  Replaces Pauli::new / parse_label / x_indices / z_indices /
 * phase (qrusty/src/lib.rs:125-181): label ^([+-]?)1?([ij]?)([IXYZ]+)$, qubit
 * q is label[len-1-q]; x bit = X or Y, z bit = Y or Z;
 * base_phase = 2*(sign=='-') + (i|j present); phase = (base_phase + #Y) mod 4.
 * Labels of more than 63 qubits are rejected with QR_ERR_SIZE (the reference
 * overflows its u64 masks there).  Host only; no GPU needed.
 * ---------------------------------------------------------------------- */
int qr_pauli_parse(const char *label, uint32_t *num_qubits, uint32_t *base_phase,
                   uint32_t *phase, uint64_t *x_mask, uint64_t *z_mask);

/* ------------------------------------------------------------------------
 * Operators.  Replaces SparsePauliOp::new / from_labels (lib.rs:354-393) plus
 * accel::rowwise::make_params (accel.rs:141-157): each term becomes
 * (x, z, c' = M[p] * c) with M = {1, -i, -1, +i}.
 *
 * phase_convention selects p:
 *   QR_PHASE_ROWWISE  p = (base_phase + #Y) mod 4   -- the rowwise/accel
 *                     builders (accel.rs:145-153) and Pauli::to_matrix_accel
 *   QR_PHASE_KRON     p = (#Y - base_phase) mod 4   -- Pauli::to_matrix, i.e.
 *                     i^base_phase times the Kronecker product (lib.rs:182-213)
 * The two agree for labels without an i/j prefix.
 *
 * Errors (messages as the reference's QrustyErr strings):
 *   QR_ERR_TERMS "SparsePauliOp::new: at least one pauli must be supplied"
 *   QR_ERR_TERMS "SparsePauliOp::new: all supplied paulis must have the same #qubits"
 *   QR_ERR_LABEL "error: malformed label"
 * (a length mismatch of labels/coeffs cannot occur: both have num_terms).
 * Host only; device tables are uploaded on first GPU use.
 * ---------------------------------------------------------------------- */
typedef struct qr_op qr_op;

enum qr_phase_convention { QR_PHASE_ROWWISE = 0, QR_PHASE_KRON = 1 };

int qr_op_create(const char *const *labels, const double *coeffs /* 2*num_terms */,
                 uint64_t num_terms, int32_t phase_convention, qr_op **out);
int qr_op_destroy(qr_op *op);

/* qr_op_create with the term tables built on the current device, for
 * operators with many terms: the labels are packed as bytes and parsed there
 * (one thread per label), the terms grouped by x mask with a stable radix sort
 * (wavefront ballots rank equal digits) and a segmented scan, and the rank
 * lists built one thread per group.  The result -- tables, errors and their
 * messages, including which failing term is reported -- is identical to
 * qr_op_create's; the device tables of the current device are already in
 * place.  At most 2^28 terms (QR_ERR_SIZE).  Synchronises the stream. */
int qr_op_create_device(const char *const *labels, const double *coeffs /* 2*num_terms */,
                        uint64_t num_terms, int32_t phase_convention, void *stream, qr_op **out);

/* The grouped tables (host copies), for inspection: gx[G] (distinct x masks,
 * ascending), gterm[G+1] (term range of each group), tz[T] and tc[2T] (z masks
 * and c' in grouped order, input order inside a group), gkind[G] (0 complex,
 * 1 real, 2 imaginary coefficients), rl_ptr[G+1] and rl[2E] ((bit, sigma)
 * rank-list entries).  Any output may be NULL; *num_rank_entries = E.
 * Host only. */
int qr_op_tables(const qr_op *op, uint64_t *gx, uint32_t *gterm, uint64_t *tz, double *tc,
                 uint8_t *gkind, uint32_t *rl_ptr, int32_t *rl, uint64_t *num_rank_entries);

/* num_qubits n, terms T, distinct x masks G, CSR nnz = 2^n * G (every row of
 * the rowwise CSR holds exactly G entries, explicit zeros kept; accel.rs:191-208).
 * nnz is 0 when 2^n * G overflows 64 bits. */
int qr_op_info(const qr_op *op, uint32_t *num_qubits, uint64_t *num_terms,
               uint64_t *num_groups, uint64_t *nnz);

/* The make_params table in input order: x[T], z[T], coeff[2T] (c').  Host only. */
int qr_op_terms(const qr_op *op, uint64_t *x, uint64_t *z, double *coeff);

/* Column order of one CSR row as the device kernels lay it out: cols[G]
 * (ascending r ^ x_g) and the group index of each position.  Host only; an
 * inspection aid for the row-ordering tables (no values are computed). */
int qr_op_row_layout(const qr_op *op, uint64_t row, uint64_t *cols, uint32_t *groups);

/* ------------------------------------------------------------------------
 * AccelMode strings.  Replaces AccelMode::try_from (lib.rs:293-331):
 * "" | Binary | Accel | Rowwise | RowwiseUnsafe | RowwiseUnsafeChunked/N |
 * Reduce | Rayon | RayonChunked/N, plus "Hip" for this library.  Every mode
 * is served by the HIP rowwise kernel; *phase_convention reports which phase
 * convention that mode uses in the reference (lib.rs:401-585).
 * Unknown strings -> QR_ERR_MODE "to_matrix_mode: unrecognized mode <s>"
 * (pyqrusty/src/lib.rs:410-411).  Host only.
 * ---------------------------------------------------------------------- */
enum qr_mode {
    QR_MODE_DEFAULT = 0, QR_MODE_BINARY = 1, QR_MODE_ACCEL = 2, QR_MODE_ROWWISE = 3,
    QR_MODE_ROWWISE_UNSAFE = 4, QR_MODE_ROWWISE_UNSAFE_CHUNKED = 5, QR_MODE_REDUCE = 6,
    QR_MODE_RAYON = 7, QR_MODE_RAYON_CHUNKED = 8, QR_MODE_HIP = 9
};
int qr_mode_parse(const char *mode, int32_t *kind, uint64_t *step, int32_t *phase_convention);

/* ------------------------------------------------------------------------
 * CSR build.  Replaces accel::rowwise::make_unsafe_vectors_chunked
 * (accel.rs:267-336) and its per-row make_row (accel.rs:171-210).
 * Row r holds, at positions [r*G, (r+1)*G), the columns r ^ x_g in ascending
 * order; the value of group g is the term-order fold (first value, then +)
 * of (-1)^popcount(r & z_t) * c'_t over the terms t with x_t = x_g.
 * Output is bit-identical to the reference's rowwise CSR.
 *   d_indptr  : N+1 uint64 (indptr[r] = r*G)
 *   d_indices : N*G uint64
 *   d_data    : N*G complex128
 * qr_csr_build_rows writes rows [row_begin, row_end) only, at positions
 * (r - row_begin)*G (no indptr); it lets callers stream matrices larger
 * than device memory.
 * ---------------------------------------------------------------------- */
int qr_csr_build(const qr_op *op, uint64_t *d_indptr, uint64_t *d_indices, double *d_data,
                 void *stream);
int qr_csr_build_rows(const qr_op *op, uint64_t row_begin, uint64_t row_end,
                      uint64_t *d_indices, double *d_data, void *stream);

/* ------------------------------------------------------------------------
 * CSR x dense vector.  Replaces accel::rowwise::spmat_dot_densevec
 * (accel.rs:338-370): y[r] = sum_k data[k] * x[indices[k]], k in row r, for
 * any CSR with uint64 indptr/indices (e.g. one made by SpMat.new_unchecked).
 * complex128 in and out; FP64 accumulation with fused multiply-adds.
 * nnz (= indptr[nrows]) sizes the team of lanes that shares a row.
 * ---------------------------------------------------------------------- */
int qr_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                  const uint64_t *d_indices, const double *d_data, const double *d_x,
                  double *d_y, void *stream);

/* ------------------------------------------------------------------------
 * Eigensolver-step operations on device vectors (complex128, interleaved).
 * Arithmetic is num-complex's: no fused multiply-add; division
 * ((a c + b d) / |q|^2, (b c - a d) / |q|^2); norm = hypot.
 *  qr_csr_diagonal : out[r] = the entry (r, r) of a CSR with ascending columns
 *                    per row, 0 where absent; r < min(nrows, ncols).  Replaces
 *                    SpMat.diagonal (pyqrusty/src/lib.rs:118-125).
 *  qr_axpby / qr_axpy / qr_ax : z = a x + b y, a x + y, a x; a and b are host
 *                    complex128 scalars (2 doubles).  Replace
 *                    qrusty::accel::axpby / axpy / ax (accel.rs:374-393).
 *  qr_precond2     : out = dx / reg(diag - e, tol), reg(v) = tol + 0i when
 *                    |v| < tol, else v (pyqrusty/src/lib.rs:436-468).
 *  qr_precond      : the same with diag read from a CSR (length min(nrows,
 *                    ncols)).
 * ---------------------------------------------------------------------- */
int qr_csr_diagonal(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                    const uint64_t *d_indices, const double *d_data, double *d_out, void *stream);
int qr_axpby(uint64_t n, const double *a, const double *d_x, const double *b, const double *d_y,
             double *d_z, void *stream);
int qr_axpy(uint64_t n, const double *a, const double *d_x, const double *d_y, double *d_z,
            void *stream);
int qr_ax(uint64_t n, const double *a, const double *d_x, double *d_z, void *stream);
int qr_precond(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
               const uint64_t *d_indices, const double *d_data, const double *d_dx,
               const double *e, double tol, double *d_out, void *stream);
int qr_precond2(uint64_t n, const double *d_diag, const double *d_dx, const double *e, double tol,
                double *d_out, void *stream);

/* a*A + b*B of two CSR matrices of the same shape with ascending columns per
 * row (pyqrusty's a_spmat_p_b_spmat, pyqrusty/src/lib.rs:499-517: sprs scale
 * and `+`).  Each output row is the merge of the two rows: a*v1 + b*v2 where
 * both hold the column, a*v1 + 0 or 0 + b*v2 otherwise; exact zeros are
 * dropped.  Two calls: qr_csr_add_count writes d_indptr_out (nrows+1) and
 * returns the output nnz on the host (it synchronises the stream); the caller
 * allocates, and qr_csr_add_fill writes indices and values. */
int qr_csr_add_count(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *d_indptr1,
                     const uint64_t *d_indices1, const double *d_data1, const double *b,
                     uint64_t nnz2, const uint64_t *d_indptr2, const uint64_t *d_indices2,
                     const double *d_data2, uint64_t *d_indptr_out, uint64_t *nnz_out,
                     void *stream);
int qr_csr_add_fill(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *d_indptr1,
                    const uint64_t *d_indices1, const double *d_data1, const double *b,
                    uint64_t nnz2, const uint64_t *d_indptr2, const uint64_t *d_indices2,
                    const double *d_data2, const uint64_t *d_indptr_out, uint64_t *d_indices_out,
                    double *d_data_out, void *stream);

/* Tolerance filters (qrusty/src/util.rs:144-171; SpMat.count_zeros and
 * SpMat.eliminate_zeros, pyqrusty/src/lib.rs:170-183).  |v| = hypot(re, im).
 *  qr_csr_count_zeros     : *count = #entries with |v| <= tol (synchronises).
 *  qr_csr_eliminate_zeros : the CSR of the entries with |v| > tol, in their
 *                           order; outputs sized for nnz entries (the upper
 *                           bound), *nnz_out = kept entries (synchronises). */
int qr_csr_count_zeros(uint64_t nnz, const double *d_data, double tol, uint64_t *count,
                       void *stream);
int qr_csr_eliminate_zeros(uint64_t nrows, uint64_t nnz, const uint64_t *d_indptr,
                           const uint64_t *d_indices, const double *d_data, double tol,
                           uint64_t *d_indptr_out, uint64_t *d_indices_out, double *d_data_out,
                           uint64_t *nnz_out, void *stream);

/* ------------------------------------------------------------------------
 * CSR files of a device matrix.  A path whose extension is "gz" is gzip
 * (qrusty/src/util.rs:316-369).  Entries stream between the device and the
 * file through pinned buffers; the calls synchronise the stream.
 *
 * Raw binary format of qrusty::rawio::write / read (qrusty/src/rawio.rs:
 * 128-179): "MI", u64 storage (0 = CSR), rows, cols, then indptr, indices and
 * data each preceded by its u64 length, in the writer's byte order (a file
 * written with the other byte order is swapped on reading).
 *  qr_csr_write_raw     : writes a CSR of rows + 1 indptr words.
 *  qr_csr_read_raw_info : rows, cols and nnz of a file (CSR storage only).
 *  qr_csr_read_raw      : reads it into caller buffers of rows + 1, nnz and
 *                         2 * nnz words (rows and nnz as the info call gave).
 *
 * MatrixMarket coordinate complex (SpMat.matrixmarket_write / _read /
 * _write_symmetric, pyqrusty/src/lib.rs:216-259; sprs io 0.11): header
 * "%%MatrixMarket matrix coordinate complex <symmetry>", "% written by sprs",
 * "rows cols count", then "row col re im" per written entry (1-based; the
 * floats as Rust's Display prints them).  QR_MM_GENERAL writes every entry;
 * the symmetric modes the lower triangle (row >= col; row > col for
 * skew-symmetric).  Reading accepts complex, real and integer fields and
 * every symmetry, expands the symmetric ones (mirror v, -v or conj v), sums
 * duplicates in file order and sorts the columns of each row.
 *  qr_mm_read_open  : parses a file into a reader holding the device CSR and
 *                     reports its shape and nnz.
 *  qr_mm_read_take  : copies the CSR into caller buffers (rows + 1, nnz,
 *                     2 * nnz words).
 *  qr_mm_read_close : frees the reader.
 * ---------------------------------------------------------------------- */
enum { QR_MM_GENERAL = 0, QR_MM_SYMMETRIC = 1, QR_MM_HERMITIAN = 2, QR_MM_SKEW_SYMMETRIC = 3 };
typedef struct qr_mm_reader qr_mm_reader;

int qr_csr_write_raw(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz,
                     const uint64_t *d_indptr, const uint64_t *d_indices, const double *d_data,
                     void *stream);
int qr_csr_read_raw_info(const char *path, uint64_t *rows, uint64_t *cols, uint64_t *nnz);
int qr_csr_read_raw(const char *path, uint64_t rows, uint64_t nnz, uint64_t *d_indptr,
                    uint64_t *d_indices, double *d_data, void *stream);
int qr_csr_write_mm(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz,
                    const uint64_t *d_indptr, const uint64_t *d_indices, const double *d_data,
                    int32_t symmetry, void *stream);
int qr_mm_read_open(const char *path, qr_mm_reader **out, uint64_t *rows, uint64_t *cols,
                    uint64_t *nnz, void *stream);
int qr_mm_read_take(const qr_mm_reader *reader, uint64_t *d_indptr, uint64_t *d_indices,
                    double *d_data, void *stream);
int qr_mm_read_close(qr_mm_reader *reader);

/* ------------------------------------------------------------------------
 * Matrix-free Pauli-sum apply (no counterpart in the reference; equal to the
 * CSR build followed by the CSR matvec):
 *   y_v[r] = sum_g D_g(r) * x_v[r ^ x_g],  D_g(r) = sum_{t in g} (-1)^popcount(r&z_t) c'_t
 * for nvec state vectors stored one after another (x_v at offset v*2^n).
 *   QR_PREC_FP64  : x, y complex128; FP64 throughout.
 *   QR_PREC_MIXED : x, y complex64; D_g(r) merged in FP64 and rounded to FP32,
 *                   products and the sum over groups in FP32 with TwoSum
 *                   compensation; <x|y> reduced in FP64.
 * d_expect (optional, may be NULL): 2*nvec doubles (device) receiving
 * <x_v | y_v> = sum_r conj(x_v[r]) y_v[r] in FP64.
 * d_y may be NULL when d_expect is given: then only <x|H|x> is produced and
 * y is never written.  y must not overlap x (QR_ERR_ARG): the apply is not in
 * place, since blocks gather x rows that other blocks' y rows would replace.
 * Each launch that computes <x|H|x> takes its own workspace for the block
 * partials from a per-device stream-ordered pool, so calls on one qr_op from
 * several threads and streams may overlap.
 * Two kernel forms give the same results within rounding (the order of the
 * sum over groups differs): one gathers partner tiles per group, the other
 * loads each partner tile into LDS once for all the groups that read it.
 * The library picks one from the operator's masks and the state size; the
 * environment variable QR_APPLY_FORM=1 (gather) or 2 (staged) forces it.
 * QR_PREC_MIXED takes tiles of 2^11 rows, or of 2^10 when the launch would
 * otherwise have fewer than 4 tiles per CU (results again equal within
 * rounding); QR_APPLY_SMALL=0 or 1 forces the larger or the smaller.
 * ---------------------------------------------------------------------- */
enum qr_precision { QR_PREC_FP64 = 0, QR_PREC_MIXED = 1 };

int qr_pauli_apply(const qr_op *op, int32_t precision, uint32_t nvec, const void *d_x,
                   void *d_y, double *d_expect, void *stream);

/* Allocation-free qr_pauli_apply.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same arguments and
 * results as qr_pauli_apply, but it calls no hipMalloc, hipFree or allocating
 * library routine:
 *  - the <x|y> block partials live in a file-scope device array of 4 MiB per
 *    device (2^19 doubles): 2 * nvec * (2^n >> L) doubles must fit, with
 *    L = 11 (QR_PREC_MIXED) or 10 (QR_PREC_FP64) -- e.g. 8 states of 24
 *    qubits, or one of 28 -- else QR_ERR_SIZE (complex64 launches of fewer
 *    than 4 tiles per CU take tiles of 2^10 rows, whose partials always fit);
 *  - the operator's device tables for this precision must already be on the
 *    device: one qr_pauli_apply call of the same operator and precision on
 *    that device uploads them (QR_ERR_ARG otherwise).  Complex64 has tables
 *    per tile size; when only the other size's are on the device, the call
 *    uses those (same results as the qr_pauli_apply that uploaded them).
 * Every call on one device shares that array, which keeps the last call's
 * contents: calls that overlap on one device (other streams or threads)
 * race, so the caller serialises them. */
int qr_pauli_apply_noalloc(const qr_op *op, int32_t precision, uint32_t nvec, const void *d_x,
                           void *d_y, double *d_expect, void *stream);

/* ------------------------------------------------------------------------
 * Batches (no counterpart in the reference, which builds and applies one
 * operator at a time: lib.rs:573-585, accel.rs:338-370).  One call builds the
 * CSR of every operator of the batch and one matvec y_k = A_k x_k each, as
 * persistent kernels that take (operator, row tile) items from atomic
 * counters, recorded in a hipGraph (replayed while the buffers and tuning stay
 * the same).  Split form: two build kernels, whose items are every
 * operator's tiles (as sub-operators of at most 2048 groups), then one matvec
 * kernel over the written CSRs.  Fused form: one build-and-apply kernel per
 * register class, plus the single-operator kernels for operators of more
 * than 6144 groups.  CSR entries are bit-identical to qr_csr_build; y_k is
 * computed in FP64 with FMAs.
 *   d_indptr[k], d_indices[k], d_data[k] : CSR outputs sized as for
 *       qr_csr_build, or NULL (all three, or the arrays themselves) to keep
 *       only y_k (fused form) -- except for operators with more than 6144
 *       distinct x masks, which need them; the split form needs d_indices
 *       and d_data of every operator, d_indptr is optional
 *   d_x[k], d_y[k] : complex128 states of 2^n_k entries (n_k <= 32)
 *   tuning : NULL = the tuned (or default) parameters
 * qr_batch_tune times every tuning candidate on the given buffers
 * (synchronising the stream), keeps and returns the fastest.
 * ---------------------------------------------------------------------- */
typedef struct qr_batch qr_batch;
typedef struct {
    uint32_t tiles_per_item;  /* row tiles per work item (1..4096)          */
    uint32_t blocks_per_cu;   /* persistent blocks per compute unit (1..8)  */
    uint32_t order;           /* 0: input order, 1: largest operator first  */
    uint32_t split;           /* 0: matvec fused into the build (reads no CSR
                                 back); 1: build pass, then a matvec pass over
                                 the written CSR (needs every CSR buffer)     */
} qr_tuning;

int qr_batch_create(const qr_op *const *ops, uint32_t nops, qr_batch **out);
int qr_batch_destroy(qr_batch *batch);
int qr_batch_run(qr_batch *batch, uint64_t *const *d_indptr, uint64_t *const *d_indices,
                 double *const *d_data, const double *const *d_x, double *const *d_y,
                 const qr_tuning *tuning, void *stream);
int qr_batch_tune(qr_batch *batch, uint64_t *const *d_indptr, uint64_t *const *d_indices,
                  double *const *d_data, const double *const *d_x, double *const *d_y,
                  qr_tuning *best, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QRHIP_H */
