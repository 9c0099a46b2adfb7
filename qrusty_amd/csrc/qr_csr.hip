// qr_csr.hip -- CSR build and CSR x vector kernels for gfx950 (CDNA4, wave64).
//
// CSR build (replaces accel::rowwise::make_row / make_unsafe_vectors_chunked,
// qrusty/src/accel.rs:171-210, 267-336).  The reference builds each row by
// emitting T (col, val) pairs, stable-sorting them by column and folding equal
// columns.  Here the work is split by *output entry*: one thread produces the
// entry of (row r, group j), where the groups are the distinct x masks sorted
// ascending and each group's terms are stored in input order.  Nothing is
// sorted on the device:
//   column   = r ^ x_j
//   position = r*G + rank_r(j), rank_r(j) = j + sum_{(b, s) in list_j, r_b=1} s
//              (the trie identity documented in qr_api.cpp: build_rank_lists)
//   value    = fold over the group's terms, in input order, of
//              (-1)^popcount(r & z_t) * c'_t  (first value, then +)
// which is exactly the reference's stable-sort-then-fold result, so indices
// and values are bit-identical.  The kernel is bound by HBM writes
// (24 B per nnz); consecutive lanes take consecutive groups of one row, so a
// wave's stores land in a few contiguous runs of that row.
//
// CSR matvec (replaces accel::rowwise::spmat_dot_densevec, accel.rs:338-370):
// a power-of-two team of lanes per row streams the row's indices/values with
// non-temporal loads, gathers x, accumulates in FP64 with FMAs and reduces the
// team with cross-lane shuffles.
#include "qr_internal.h"

namespace qr {

typedef double dvec2 __attribute__((ext_vector_type(2)));

constexpr int kBuildThreads = 256;
constexpr int kBuildItems = 8;  // entries per thread per block

template <bool kNarrow>
__device__ __forceinline__ uint32_t parity_and(uint64_t a, uint64_t b) {
    if (kNarrow) return __popc(uint32_t(a) & uint32_t(b)) & 1u;
    return __popcll(a & b) & 1u;
}

template <bool kNarrow>
__global__ __launch_bounds__(kBuildThreads) void k_csr_build(
    uint64_t row_begin, uint64_t nrows, uint32_t G, uint32_t step_q, uint32_t step_r,
    const uint64_t *__restrict__ gx, const uint32_t *__restrict__ gterm,
    const uint64_t *__restrict__ tz, const double2 *__restrict__ tc,
    const uint32_t *__restrict__ rl_ptr, const RankEntry *__restrict__ rl,
    uint64_t *__restrict__ indices, double2 *__restrict__ data) {
    const uint64_t total = nrows * uint64_t(G);
    uint64_t item = uint64_t(blockIdx.x) * (kBuildThreads * kBuildItems) + threadIdx.x;
    if (item >= total) return;
    uint64_t lo = item / G;
    uint32_t j = uint32_t(item - lo * G);
#pragma unroll 1
    for (int k = 0; k < kBuildItems; ++k) {
        const uint64_t r = row_begin + lo;
        // position of group j in row r
        int64_t rank = int64_t(j);
        const uint32_t e1 = rl_ptr[j + 1];
        for (uint32_t e = rl_ptr[j]; e < e1; ++e) {
            const RankEntry en = rl[e];
            if ((r >> en.bit) & 1) rank += en.sigma;
        }
        // term-order fold of the group's signed coefficients
        const uint32_t t0 = gterm[j], t1 = gterm[j + 1];
        double2 c = tc[t0];
        if (parity_and<kNarrow>(r, tz[t0])) { c.x = -c.x; c.y = -c.y; }
        double2 v = c;
        for (uint32_t t = t0 + 1; t < t1; ++t) {
            c = tc[t];
            if (parity_and<kNarrow>(r, tz[t])) { c.x = -c.x; c.y = -c.y; }
            v.x = v.x + c.x;
            v.y = v.y + c.y;
        }
        const uint64_t pos = lo * G + uint64_t(rank);
        __builtin_nontemporal_store(r ^ gx[j], indices + pos);
        dvec2 vv = {v.x, v.y};
        __builtin_nontemporal_store(vv, reinterpret_cast<dvec2 *>(data + pos));

        item += kBuildThreads;
        if (item >= total) break;
        j += step_r;
        lo += step_q;
        if (j >= G) { j -= G; ++lo; }
    }
}

__global__ __launch_bounds__(256) void k_indptr(uint64_t row_begin, uint64_t nrows, uint64_t G,
                                                uint64_t *__restrict__ indptr) {
    uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i <= nrows) indptr[i] = (row_begin + i) * G;
}

int launch_csr_build(const qr_op *op, const DeviceTables *t, uint64_t row_begin, uint64_t row_end,
                     uint64_t *d_indptr, uint64_t *d_indices, double *d_data, hipStream_t s) {
    const uint64_t nrows = row_end - row_begin;
    const uint32_t G = uint32_t(op->G);
    if (op->G > 0xFFFFFFFFull) return fail(QR_ERR_SIZE, "too many groups");
    if (d_indptr) {
        uint64_t nb = (nrows + 1 + 255) / 256;
        hipLaunchKernelGGL(k_indptr, dim3(unsigned(nb)), dim3(256), 0, s, row_begin, nrows,
                           uint64_t(G), d_indptr);
        QR_HIP_CHECK(hipGetLastError());
    }
    const uint64_t total = nrows * uint64_t(G);
    const uint64_t per_block = uint64_t(kBuildThreads) * kBuildItems;
    const uint64_t nblocks = (total + per_block - 1) / per_block;
    if (nblocks > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR build grid too large");
    const uint32_t step_q = kBuildThreads / G, step_r = kBuildThreads % G;
    const bool narrow = op->nq <= 32;
    auto kern = narrow ? k_csr_build<true> : k_csr_build<false>;
    hipLaunchKernelGGL(kern, dim3(unsigned(nblocks)), dim3(kBuildThreads), 0, s, row_begin, nrows,
                       G, step_q, step_r, t->gx, t->gterm, t->tz, t->tc, t->rl_ptr, t->rl,
                       d_indices, reinterpret_cast<double2 *>(d_data));
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

// ---------------------------------------------------------------------------
// CSR x vector
// ---------------------------------------------------------------------------
// Entries outside [0, nnz) and columns >= ncols are skipped, so a malformed
// CSR handed in through SpMat.new_unchecked cannot read out of bounds.
template <int LPR>
__global__ __launch_bounds__(256) void k_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz,
                                                    const uint64_t *__restrict__ indptr,
                                                    const uint64_t *__restrict__ indices,
                                                    const double2 *__restrict__ data,
                                                    const double2 *__restrict__ x,
                                                    double2 *__restrict__ y) {
    const uint64_t gid = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    const uint64_t row = gid / LPR;
    const int lane = int(threadIdx.x % LPR);
    double2 acc = make_double2(0.0, 0.0);
    if (row < nrows) {
        const uint64_t k0 = min(indptr[row], nnz), k1 = min(indptr[row + 1], nnz);
        uint64_t k = k0 + lane;
        // four independent loads in flight per lane
        for (; k + 3 * LPR < k1; k += 4 * LPR) {
            uint64_t c[4];
            double2 a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                c[u] = __builtin_nontemporal_load(indices + k + u * LPR);
                const dvec2 av = __builtin_nontemporal_load(
                    reinterpret_cast<const dvec2 *>(data + k + u * LPR));
                a[u] = make_double2(av.x, av.y);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double2 b = make_double2(0.0, 0.0);
                if (c[u] < ncols) b = x[c[u]];
                acc.x = fma(a[u].x, b.x, acc.x);
                acc.x = fma(-a[u].y, b.y, acc.x);
                acc.y = fma(a[u].x, b.y, acc.y);
                acc.y = fma(a[u].y, b.x, acc.y);
            }
        }
        for (; k < k1; k += LPR) {
            const uint64_t c = __builtin_nontemporal_load(indices + k);
            const dvec2 av = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(data + k));
            const double2 a = make_double2(av.x, av.y);
            double2 b = make_double2(0.0, 0.0);
            if (c < ncols) b = x[c];
            acc.x = fma(a.x, b.x, acc.x);
            acc.x = fma(-a.y, b.y, acc.x);
            acc.y = fma(a.x, b.y, acc.y);
            acc.y = fma(a.y, b.x, acc.y);
        }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off /= 2) {
        acc.x += __shfl_xor(acc.x, off, LPR);
        acc.y += __shfl_xor(acc.y, off, LPR);
    }
    if (row < nrows && lane == 0) y[row] = acc;
}

int launch_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                      const uint64_t *d_indices, const double *d_data, const double *d_x,
                      double *d_y, hipStream_t s) {
    const uint64_t avg = nrows ? nnz / nrows : 0;
    int lpr = 1;
    while (lpr < 64 && uint64_t(lpr) * 2 <= avg) lpr *= 2;  // ~2+ entries per lane
    const uint64_t threads = nrows * uint64_t(lpr);
    const uint64_t nblocks = (threads + 255) / 256;
    if (nblocks > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR matvec grid too large");
    auto x = reinterpret_cast<const double2 *>(d_x);
    auto data = reinterpret_cast<const double2 *>(d_data);
    auto y = reinterpret_cast<double2 *>(d_y);
    const dim3 grid(static_cast<unsigned>(nblocks)), block(256);
    switch (lpr) {
        case 1: hipLaunchKernelGGL(k_csr_matvec<1>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
        case 2: hipLaunchKernelGGL(k_csr_matvec<2>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
        case 4: hipLaunchKernelGGL(k_csr_matvec<4>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
        case 8: hipLaunchKernelGGL(k_csr_matvec<8>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
        case 16: hipLaunchKernelGGL(k_csr_matvec<16>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
        case 32: hipLaunchKernelGGL(k_csr_matvec<32>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
        default: hipLaunchKernelGGL(k_csr_matvec<64>, grid, block, 0, s, nrows, ncols, nnz, d_indptr, d_indices, data, x, y); break;
    }
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

}  // namespace qr
