// qr_internal.h -- shared declarations of libqrhip (host tables, launchers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qrhip.h"

struct qr_op;

// a parsed MatrixMarket file: the CSR in device memory (qr_mm_read_open)
struct qr_mm_reader {
    uint64_t rows = 0, cols = 0, nnz = 0;
    uint64_t *indptr = nullptr, *indices = nullptr;  // device
    double *data = nullptr;                          // device, re/im pairs
    ~qr_mm_reader() {
        if (indptr) (void)hipFree(indptr);
        if (indices) (void)hipFree(indices);
        if (data) (void)hipFree(data);
    }
};

namespace qr {

// Sets the thread's last error and returns `code`.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define QR_HIP_CHECK(call)                                      \
    do {                                                        \
        hipError_t qr_e_ = (call);                              \
        if (qr_e_ != hipSuccess) return ::qr::hip_fail(qr_e_, #call); \
    } while (0)

constexpr int kMaxDevices = 16;
// heavy groups the CSR build folds per row in a pass of their own: molecular
// Hamiltonians (the reference's H8..H12) have 17-111 groups of 22-301 terms
// next to groups of at most 6 (as many as the tile's LDS holds are folded
// that way, heavy_fit in qr_tile.h)
constexpr uint32_t kMaxHeavy = 128;
// Tail table of an operator with heavy groups: terms 2, 3, ... of every group
// transposed, slot u * G + g = term 2 + u of group g, for u < ntail <=
// kMaxTail and at most kMaxTailSlots slots (24 B each: at most 96 MiB per
// device; the builder reads the terms past the table from the grouped table)
constexpr uint32_t kMaxTail = 16;
constexpr uint64_t kMaxTailSlots = uint64_t(1) << 22;
constexpr uint32_t kSubMaxG = 2048;  // groups of one CSR-build sub-operator
// CSR matvecs load the CSR stream nontemporally from this many entries per
// row on (measured crossover; qr_csr.hip)
constexpr uint64_t kMvNontemporalMinRow = 256;

// Row-ordering entry: while building row r, a group whose sorted-x position is j
// moves by `sigma` positions when bit `bit` of r is set (see build_rank_lists).
struct RankEntry {
    int32_t bit;
    int32_t sigma;
};

// One (sub)group of the matrix-free apply: x mask, its pattern range split by
// class -- [p0, ps) z bits among the thread-index bits only, [ps, pk)
// straddling, [pk, p1) row-index bits only -- and its row in the chunk's
// per-tile table (-1 if it has no row-index patterns).  Pattern indices are
// relative to the chunk.  32 bytes, read with one scalar load.
struct ApplyGroup {
    uint64_t m;
    uint32_t p0, ps, pk, p1;
    int32_t u;
    uint32_t flags;  // kGroupReal
};
constexpr uint32_t kGroupReal = 1;  // every coefficient of the group is real

// Groups whose patterns [p0, p1) and table rows [u0, u1) are staged in LDS
// together.  Inside a chunk the groups are ordered by class -- bit 0: the mask
// reaches outside the tile ("far"), bit 1: D varies by row (row-bit or
// straddling patterns), bit 2: complex coefficients, bit 3 (paired-row layout
// only): row bit 0 of the mask set, so that each row's partner is the other
// half of the gathered pair -- and class c holds the groups [cg[c], cg[c+1]).
// For the partner-staged kernel the groups are ordered by partner tile
// instead, in runs [r0, r1) (ApplyRun).
constexpr int kApplyClasses = 16;
struct ApplyChunk {
    uint32_t cg[kApplyClasses + 1];
    uint32_t p0, p1, u0, u1;
    uint32_t r0, r1;
    uint32_t pad;
};

// Partner-staged apply: the groups of a chunk whose masks share the bits
// above the tile (`hi`, a row offset) gather from the same partner tile,
// which is loaded into LDS once for all of them.  Inside a run the groups are
// ordered by class -- bit 0: D varies by row, bit 1: complex coefficients --
// and class c holds the groups [g[c], g[c+1]).
constexpr int kRunClasses = 4;
struct ApplyRun {
    uint64_t hi;
    uint32_t g[kRunClasses + 1];
    uint32_t pad;
};

// Device copies of an operator's tables (one set per device that used it).
struct DeviceTables {
    bool ready = false;
    uint64_t *gx = nullptr;      // [G]   distinct x masks, ascending
    uint32_t *gterm = nullptr;   // [G+1] term range of each group
    uint64_t *tz = nullptr;      // [T]   z masks, grouped (term order inside a group)
    double2 *tc = nullptr;       // [T]   c', grouped
    uint32_t *rl_ptr = nullptr;  // [G+1]
    RankEntry *rl = nullptr;     // [E]
    uint8_t *gkind = nullptr;    // [G]   0 complex, 1 real, 2 imaginary coefficients
    uint32_t *heavy = nullptr;   // [nheavy] heavy groups (qr_op::heavy)
    uint32_t nheavy = 0;
    uint64_t *ttz = nullptr;     // [ntail * G] tail table (qr_op::ttz, ttc)
    double2 *ttc = nullptr;
    uint32_t ntail = 0;
    uint32_t *leaf_heavy = nullptr;  // qr_op::leaf_heavy
    // matrix-free apply tables, one set per tile kind and form (slot =
    // 2 * kind + staged; kind 0: FP64, 1: complex64, 2: small complex64 tiles)
    struct Patterns {
        int L = -1;
        uint32_t P = 0, ngrp = 0, nchunk = 0, nrun = 0, max_p = 0, max_u = 0;
        uint32_t nfar = 0;  // groups whose mask reaches outside the tile
        ApplyGroup *grp = nullptr;    // [ngrp]
        ApplyChunk *chunk = nullptr;  // [nchunk]
        ApplyRun *run = nullptr;      // [nrun] (staged form only)
        uint32_t *ug = nullptr;     // [U]   group (index into grp) of each table row
        uint32_t *pzl = nullptr;    // [P]   z & (2^L - 1) of the pattern
        uint32_t *ptptr = nullptr;  // [P+1] term range of each pattern
        uint64_t *ptz = nullptr;    // [T]   z masks in pattern order
        double2 *ptc = nullptr;     // [T]   c' in pattern order
    } pat[6];
};

// Host build of the apply tables for tile size L, tb thread-index bits, at
// most pcap patterns and ucap table rows per chunk (see qr_apply.hip).  A
// group with more than pcap patterns is split into several ApplyGroups with
// the same mask.  `staged` orders each chunk by partner tile (ApplyRun)
// instead of by class.
struct HostPatterns {
    std::vector<ApplyGroup> grp;
    std::vector<ApplyChunk> chunk;
    std::vector<uint32_t> ug, pzl, ptptr;
    std::vector<uint64_t> ptz;
    std::vector<double> ptc;  // 2 per term
    std::vector<ApplyRun> run;  // staged form only
    uint32_t max_p = 0, max_u = 0, nfar = 0;
};
// Paired-row layout of the mixed gather form: thread t owns the row pairs
// (2t, 2t+1) + 2^(tb+1) j, so row-bit 0 counts with the row index (bit 0 of
// k) and row-bits 1..tb are the thread bits.  The patterns then hold z in
// that order: bits 1..tb moved down to 0..tb-1, bit 0 up to bit tb.
inline uint64_t pair_row_bits(uint64_t zl, int tb) {
    const uint64_t tmask = (uint64_t(1) << tb) - 1;
    return ((zl >> 1) & tmask) | ((zl & 1) << tb) | (zl & ~((tmask << 1) | 1));
}
void build_patterns(const qr_op *op, int L, int tb, uint32_t pcap, uint32_t ucap, bool staged,
                    bool pairs, HostPatterns &out);
// Whether slot `slot` holds the apply tables for tile size L.
bool tables_ready(const qr_op *op, const DeviceTables *t, int slot, int L);
// With `noalloc` the tables must already be on the device (QR_ERR_ARG if not):
// nothing is allocated or uploaded.
int device_patterns(const qr_op *op, DeviceTables *t, int slot, int L, int tb, uint32_t pcap,
                    uint32_t ucap, bool staged, bool pairs, bool noalloc = false);

}  // namespace qr

struct qr_op {
    uint32_t nq = 0;
    uint64_t T = 0, G = 0;
    // make_params table, input order
    std::vector<uint64_t> x, z;
    std::vector<double> cp;  // 2*T
    // grouped tables (groups ascending by x; terms of a group in input order)
    std::vector<uint64_t> gx;
    std::vector<uint32_t> gterm;
    std::vector<uint64_t> tz;
    std::vector<double> tc;
    std::vector<uint8_t> gkind;
    std::vector<uint32_t> rl_ptr;
    std::vector<qr::RankEntry> rl;
    // groups with many more terms than the others (at most kMaxHeavy,
    // ascending): the CSR build folds them once per tile row, not per lane
    std::vector<uint32_t> heavy;
    // G > kSubMaxG: the mask trie cut into subtrees of at most kSubMaxG
    // groups [s0, s1) at depth `depth`; the CSR build runs one sub-operator
    // per subtree (qr_tile.h TileOp::gbase)
    struct SubLeaf {
        uint32_t s0, s1, depth;
        uint32_t hoff, hcnt;  // its heavy groups: leaf_heavy[hoff, hoff + hcnt)
    };
    std::vector<SubLeaf> leaves;
    std::vector<uint32_t> leaf_heavy;  // heavy groups of each leaf, leaf-local
    // operators with heavy groups: the terms past each group's second,
    // transposed (slot u * G + g holds term gterm[g] + 2 + u, zero where the
    // group is shorter), so that the lanes of a batch -- one group each --
    // read one contiguous run per term position instead of one line apiece
    uint32_t ntail = 0;
    std::vector<uint64_t> ttz;
    std::vector<double> ttc;  // 2 * ntail * G
    // device state
    mutable std::mutex mu;
    mutable qr::DeviceTables dev[qr::kMaxDevices];
};

namespace qr {

// An operator whose term tables are built on the current device
// (qr_group.hip): the same tables as qr_op_create's host path, and its device
// tables for this device already in place.  Synchronises the stream.
int group_terms_device(const char *const *labels, const double *coeffs, uint64_t T, int32_t conv,
                       hipStream_t s, qr_op **out);

// Sub-operators, heavy groups and tail table of an op whose grouped tables
// are set.
void finish_op(qr_op *op);
// Uploads the op's tail table (if it has one) into t.
int upload_tail(const qr_op *op, DeviceTables &t);
// Frees one device's tables of an op.
void free_tables(DeviceTables &t);

// Uploads the op's tables to the current device if needed; returns the tables.
int device_tables(const qr_op *op, DeviceTables **out, bool noalloc = false);

// Kernel launchers (qr_csr.hip, qr_apply.hip).
int launch_csr_build(const qr_op *op, const DeviceTables *t, uint64_t row_begin,
                     uint64_t row_end, uint64_t *d_indptr, uint64_t *d_indices,
                     double *d_data, hipStream_t s);
int launch_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                      const uint64_t *d_indices, const double *d_data, const double *d_x,
                      double *d_y, hipStream_t s);
int launch_csr_diagonal(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *indptr,
                        const uint64_t *indices, const double *data, double *out, hipStream_t s);
int launch_axpby(uint64_t n, int mode, const double *a, const double *x, const double *b,
                 const double *y, double *z, hipStream_t s);
int launch_precond(uint64_t n, const double *diag, uint64_t nnz, const uint64_t *indptr,
                   const uint64_t *indices, const double *data, const double *dx, const double *e,
                   double tol, double *out, hipStream_t s);
int launch_spadd_count(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *ip1,
                       const uint64_t *ix1, const double *d1, const double *b, uint64_t nnz2,
                       const uint64_t *ip2, const uint64_t *ix2, const double *d2, uint64_t *ipo,
                       uint64_t *nnz_out, hipStream_t s);
int launch_spadd_fill(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *ip1,
                      const uint64_t *ix1, const double *d1, const double *b, uint64_t nnz2,
                      const uint64_t *ip2, const uint64_t *ix2, const double *d2,
                      const uint64_t *ipo, uint64_t *ixo, double *dout, hipStream_t s);
int launch_count_zeros(uint64_t nnz, const double *d, double tol, uint64_t *count, hipStream_t s);
int launch_eliminate_zeros(uint64_t nrows, uint64_t nnz, const uint64_t *ip, const uint64_t *ix,
                           const double *d, double tol, uint64_t *ipo, uint64_t *ixo, double *dout,
                           uint64_t *nnz_out, hipStream_t s);
int io_write_raw(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz, const uint64_t *ip,
                 const uint64_t *ix, const double *d, hipStream_t s);
int io_read_raw_info(const char *path, uint64_t *rows, uint64_t *cols, uint64_t *nnz);
int io_read_raw(const char *path, uint64_t rows, uint64_t nnz, uint64_t *ip, uint64_t *ix, double *d,
                hipStream_t s);
int io_write_mm(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz, const uint64_t *ip,
                const uint64_t *ix, const double *d, int sym, hipStream_t s);
int io_read_mm(const char *path, qr_mm_reader **out, hipStream_t s);
// noalloc: no device allocation -- the tables must be on the device already
// and the <x|y> block partials go to the file-scope scratch array (qr_apply.hip)
int launch_pauli_apply(const qr_op *op, DeviceTables *t, int32_t precision, uint32_t nvec,
                       const void *d_x, void *d_y, double *d_expect, hipStream_t s,
                       bool noalloc = false);

}  // namespace qr
