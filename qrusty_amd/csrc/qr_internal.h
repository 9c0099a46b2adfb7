// qr_internal.h -- shared declarations of libqrhip (host tables, launchers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qrhip.h"

struct qr_op;

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

// Row-ordering entry: while building row r, a group whose sorted-x position is j
// moves by `sigma` positions when bit `bit` of r is set (see build_rank_lists).
struct RankEntry {
    int32_t bit;
    int32_t sigma;
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
    double *partials = nullptr;  // workspace for <x|y> block partials
    size_t partials_cap = 0;     // in doubles
    // matrix-free apply: terms of each group merged by their z bits inside a
    // tile of 2^L rows ("patterns"), one table set per tile size
    struct Patterns {
        int L = -1;
        uint32_t P = 0, nU = 0;
        uint32_t *gpat = nullptr;   // [G+1] pattern range of each group
        uint32_t *gcls = nullptr;   // [2G]  first straddling / first row-index pattern
        int32_t *gU = nullptr;      // [G]   row of the group in the per-tile k table, or -1
        uint32_t *ug = nullptr;     // [nU]  groups that have row-index patterns
        uint32_t *pzl = nullptr;    // [P]   z & (2^L - 1) of the pattern
        uint32_t *ptptr = nullptr;  // [P+1] term range of each pattern
        uint64_t *ptz = nullptr;    // [T]   z masks in pattern order
        double2 *ptc = nullptr;     // [T]   c' in pattern order
    } pat[2];
};

// Host build of the pattern tables for tile size L and tb thread-index bits
// (see qr_apply.hip).  Patterns of a group are ordered by class: z bits only
// among the thread-index bits, then straddling, then only row-index bits.
struct HostPatterns {
    std::vector<uint32_t> gpat, gcls, ug, pzl, ptptr;
    std::vector<int32_t> gU;
    std::vector<uint64_t> ptz;
    std::vector<double> ptc;  // 2 per term
};
void build_patterns(const qr_op *op, int L, int tb, HostPatterns &out);
int device_patterns(const qr_op *op, DeviceTables *t, int slot, int L, int tb);

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
    // device state
    mutable std::mutex mu;
    mutable qr::DeviceTables dev[qr::kMaxDevices];
};

namespace qr {

// Uploads the op's tables to the current device if needed; returns the tables.
int device_tables(const qr_op *op, DeviceTables **out);

// Kernel launchers (qr_csr.hip, qr_apply.hip).
int launch_csr_build(const qr_op *op, const DeviceTables *t, uint64_t row_begin,
                     uint64_t row_end, uint64_t *d_indptr, uint64_t *d_indices,
                     double *d_data, hipStream_t s);
int launch_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                      const uint64_t *d_indices, const double *d_data, const double *d_x,
                      double *d_y, hipStream_t s);
int launch_pauli_apply(const qr_op *op, DeviceTables *t, int32_t precision, uint32_t nvec,
                       const void *d_x, void *d_y, double *d_expect, hipStream_t s);

}  // namespace qr
