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
// and values are bit-identical.  The kernel is bound by HBM writes (24 B per
// nnz).  Rows are staged in LDS by the tile builder (qr_tile.h) and copied out
// coalesced; an operator with more than kSubMaxG groups is built as
// sub-operators, subtrees of the mask trie that fill a contiguous segment of
// every row (qr_tile.h TileOp::gbase).
//
// CSR matvec (replaces accel::rowwise::spmat_dot_densevec, accel.rs:338-370):
// a power-of-two team of lanes per row streams the row's indices/values with
// non-temporal loads, gathers x, accumulates in FP64 with FMAs and reduces the
// team with cross-lane shuffles.
#include "qr_internal.h"
#include "qr_mv.h"
#include "qr_tile.h"

#include <algorithm>
#include <mutex>
#include <type_traits>

namespace qr {

// LDS-staged build: one block per tile of the shared tile builder
// (qr_tile.h), without the fused matvec; the heavy-group instances of
// fold_tiles run 2^op.fold_log2 consecutive tiles per block, folding the
// heavy groups once for all of them.
template <bool kNarrow, int kChunks, int kMode>
__global__ __launch_bounds__(kStageThreads) void k_csr_build_staged(TileOp op) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if constexpr (fold_tiles<kChunks, kMode>()) {
        const uint32_t tile_log2 = 10u - op.jw_log2 + 4u;
        const uint32_t k = 1u << op.fold_log2;
        for (uint32_t s = 0; s < k; ++s) {
            const uint64_t tile = uint64_t(blockIdx.x) * k + s;
            if ((((op.row_begin >> tile_log2) + tile) << tile_log2) >= op.row_end) break;
            build_tile_staged<kNarrow, kChunks, false, kMode>(op, tile, smem_raw, s);
        }
    } else {
        build_tile_staged<kNarrow, kChunks, false, kMode>(op, blockIdx.x, smem_raw);
    }
}

// Sub-operators of at most 32 qubits (at most 2048 groups, a 48 KB stage):
// held to 64 VGPRs, i.e. 8 waves per SIMD, so that two 1024-thread blocks
// share a CU
template <bool kNarrow, int kChunks, int kMode>
__global__ __launch_bounds__(kStageThreads, 8) void k_csr_build_sub(TileOp op) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    build_tile_staged<kNarrow, kChunks, false, kMode>(op, blockIdx.x, smem_raw);
}

__global__ __launch_bounds__(256) void k_indptr(uint64_t row_begin, uint64_t nrows, uint64_t G,
                                                uint64_t *__restrict__ indptr) {
    uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i <= nrows) indptr[i] = (row_begin + i) * G;
}

int launch_csr_build(const qr_op *op, const DeviceTables *t, uint64_t row_begin, uint64_t row_end,
                     uint64_t *d_indptr, uint64_t *d_indices, double *d_data, hipStream_t s) {
    const uint64_t nrows = row_end - row_begin;
    if (op->G > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "too many groups");
    const uint32_t G = uint32_t(op->G);
    if (d_indptr) {
        uint64_t nb = (nrows + 1 + 255) / 256;
        hipLaunchKernelGGL(k_indptr, dim3(unsigned(nb)), dim3(256), 0, s, row_begin, nrows,
                           uint64_t(G), d_indptr);
        QR_HIP_CHECK(hipGetLastError());
    }
    if (nrows == 0) return QR_OK;
    const bool narrow = op->nq <= 32;
    if (!op->leaves.empty()) {
        // one sub-operator per subtree of the mask trie, each <= kSubMaxG
        // groups with one row per batch (JW = 1024): the fast 1- and 2-chunk
        // kernels, each writing its contiguous segment of every row
        const uint32_t jw = 10, tl2 = 4;
        const uint64_t nt = ((row_end - 1) >> tl2) - (row_begin >> tl2) + 1;
        if (nt > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR build grid too large");
        for (size_t k = 0; k < op->leaves.size(); ++k) {
            const auto &lf = op->leaves[k];
            const uint32_t g = lf.s1 - lf.s0, nch = (g + 1023) >> 10;
            uint32_t dbuf = 0;
            const uint32_t nh = heavy_fit(lf.hcnt, jw);
            // (sub-operators keep one stage: the second would cost them
            // tiles per heavy fold, H10 14.12 -> 14.38 ms)
            const size_t sb = stage_bytes(g, jw, &dbuf);
            const uint32_t fl = fold_tiles_log2(true, sb, heavy_bytes(nh, jw), nt);
            const size_t shmem = sb + (heavy_bytes(nh, jw) << fl);
            TileOp top{t->gx + lf.s0, t->gterm + lf.s0, t->tz, t->tc, t->rl_ptr + lf.s0, t->rl,
                       d_indices, reinterpret_cast<double2 *>(d_data), nullptr, nullptr, nullptr,
                       row_begin, row_end, g, jw, nch, dbuf,
                       nh ? t->leaf_heavy + lf.hoff : nullptr, nh, lf.s0, lf.depth, G,
                       t->ntail ? t->ttz + lf.s0 : nullptr, t->ntail ? t->ttc + lf.s0 : nullptr, t->ntail, fl};
            // (wide masks, and heavy groups, do not fit 64 VGPRs without
            // spills: one block per CU)
            auto kst = nh ? (narrow ? (nch <= 1 ? k_csr_build_staged<true, 1, kTileSubHeavy>
                                                : k_csr_build_staged<true, 2, kTileSubHeavy>)
                                    : (nch <= 1 ? k_csr_build_staged<false, 1, kTileSubHeavy>
                                                : k_csr_build_staged<false, 2, kTileSubHeavy>))
                          : (narrow ? (nch <= 1 ? k_csr_build_sub<true, 1, kTileSub>
                                                : k_csr_build_sub<true, 2, kTileSub>)
                                    : (nch <= 1 ? k_csr_build_staged<false, 1, kTileSub>
                                                : k_csr_build_staged<false, 2, kTileSub>));
            QR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kst),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
            hipLaunchKernelGGL(kst, dim3(unsigned(((nt - 1) >> fl) + 1)), dim3(kStageThreads), shmem, s, top);
            QR_HIP_CHECK(hipGetLastError());
        }
        return QR_OK;
    }
    // G <= kSubMaxG: one or two group chunks per thread.  Operators of 129 to
    // 1024 groups run two chunks on half the lanes per row (twice the rows per
    // batch): measured 6-14% faster at G = 150, 417, 690, 880 and 990; one
    // chunk stays faster at G = 25 and 60, and four chunks on 512 lanes were
    // 22% slower than two on 1024 at G = 1996 (one block per CU instead of two).
    uint32_t jw = stage_jw_log2(G);
    if (jw == 8 || jw == 9 || (jw == 10 && G <= 1024)) jw -= 1;
    const uint32_t nch = (G + (1u << jw) - 1) >> jw;
    const uint32_t tl2 = 10 - jw + 4;
    const uint64_t nt = ((row_end - 1) >> tl2) - (row_begin >> tl2) + 1;
    if (nt > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR build grid too large");
    uint32_t dbuf = 0;
    const uint32_t nh = heavy_fit(t->nheavy, jw);
    const size_t sb = nh ? heavy_stage(nch >= 2, G, jw, heavy_bytes(nh, jw), &dbuf) : stage_bytes(G, jw, &dbuf);
    const uint32_t fl = fold_tiles_log2(nch >= 2, sb, heavy_bytes(nh, jw), nt);
    const size_t shmem = sb + (heavy_bytes(nh, jw) << fl);
    TileOp top{t->gx, t->gterm, t->tz, t->tc, t->rl_ptr, t->rl, d_indices,
               reinterpret_cast<double2 *>(d_data), nullptr, nullptr, nullptr,
               row_begin, row_end, G, jw, nch, dbuf, t->heavy, nh, 0, 0, G, t->ttz, t->ttc, t->ntail, fl};
    auto pick = [&](auto narrow_t, auto heavy_t) {
        constexpr bool kN = decltype(narrow_t)::value;
        constexpr int kH = decltype(heavy_t)::value ? kTileHeavy : kTilePlain;
        return nch <= 1 ? k_csr_build_staged<kN, 1, kH> : k_csr_build_staged<kN, 2, kH>;
    };
    using T = std::true_type;
    using F = std::false_type;
    auto kst = narrow ? (nh ? pick(T{}, T{}) : pick(T{}, F{})) : (nh ? pick(F{}, T{}) : pick(F{}, F{}));
    QR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kst),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    hipLaunchKernelGGL(kst, dim3(unsigned(((nt - 1) >> fl) + 1)), dim3(kStageThreads), shmem, s, top);
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

// ---------------------------------------------------------------------------
// CSR x vector
// ---------------------------------------------------------------------------
// Entries outside [0, nnz) and columns >= ncols are skipped, so a malformed
// CSR handed in through SpMat.new_unchecked cannot read out of bounds.
//
// LPR lanes share a row, so a wave covers 64/LPR consecutive rows.  For the
// Pauli CSR, rows that differ only in their low bits gather x at columns that
// differ only in those bits, i.e. from the same cache lines, so a wave's
// gathers coalesce.  Blocks are handed to the 8 XCDs round-robin; the block
// index is remapped so that each XCD works on one contiguous range of rows
// and its L2 sees the neighbouring rows' x lines.
// CSR stream loads: nontemporal (kNT) for long rows, plain for short ones.
// Measured (MI355X, complex128): nontemporal wins from about 256 entries per
// row (400: 5%, 1000: 7%, 2000: 2.4%) and loses below (150: 3%, 60: 8%,
// 25 -- the 24-qubit Heisenberg CSR: 13%).
template <bool kNT, typename T>
__device__ __forceinline__ T mv_load(const T *p) {
    if constexpr (kNT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <int LPR, bool kNT>
__global__ __launch_bounds__(256) void k_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz,
                                                    uint32_t nblocks_xcd,
                                                    const uint64_t *__restrict__ indptr,
                                                    const uint64_t *__restrict__ indices,
                                                    const double2 *__restrict__ data,
                                                    const double2 *__restrict__ x,
                                                    double2 *__restrict__ y) {
    uint32_t b = blockIdx.x;
    if (nblocks_xcd) b = (b % 8) * nblocks_xcd + b / 8;
    const uint64_t gid = uint64_t(b) * 256 + threadIdx.x;
    const uint64_t row = gid / LPR;
    const int lane = int(threadIdx.x % LPR);
    double2 acc = make_double2(0.0, 0.0);
    if (row < nrows) {
        const uint64_t k0 = min(indptr[row], nnz), k1 = min(indptr[row + 1], nnz);
        uint64_t k = k0 + lane;
        // eight independent entries in flight per lane (measured best of 2, 4, 8);
        // full chunks of U * LPR entries for all lanes of the row together, so
        // that no lane of a row loops while another takes the tail
        constexpr int U = 8;
        for (; k - lane + U * LPR <= k1; k += U * LPR) {
            uint64_t c[U];
            double2 a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = mv_load<kNT>(indices + k + u * LPR);
                const dvec2 av = mv_load<kNT>(reinterpret_cast<const dvec2 *>(data + k + u * LPR));
                a[u] = make_double2(av.x, av.y);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                double2 bx = make_double2(0.0, 0.0);
                if (c[u] < ncols) bx = x[c[u]];
                acc.x = fma(a[u].x, bx.x, acc.x);
                acc.x = fma(-a[u].y, bx.y, acc.x);
                acc.y = fma(a[u].x, bx.y, acc.y);
                acc.y = fma(a[u].y, bx.x, acc.y);
            }
        }
        if (k < k1) {
            // the fewer than U entries left (all of a short row): predicated,
            // with every load issued before the first is used
            uint64_t c[U];
            double2 a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = ~uint64_t(0);
                a[u] = make_double2(0.0, 0.0);
                if (k + u * LPR < k1) {
                    c[u] = mv_load<kNT>(indices + k + u * LPR);
                    const dvec2 av = mv_load<kNT>(reinterpret_cast<const dvec2 *>(data + k + u * LPR));
                    a[u] = make_double2(av.x, av.y);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (c[u] < ncols) {
                    const double2 bx = x[c[u]];
                    acc.x = fma(a[u].x, bx.x, acc.x);
                    acc.x = fma(-a[u].y, bx.y, acc.x);
                    acc.y = fma(a[u].x, bx.y, acc.y);
                    acc.y = fma(a[u].y, bx.x, acc.y);
                }
            }
        }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off /= 2) {
        acc.x += __shfl_xor(acc.x, off, LPR);
        acc.y += __shfl_xor(acc.y, off, LPR);
    }
    if (row < nrows && lane == 0) y[row] = acc;
}

// Rows of 16..32 entries on average (8 lanes per row): each 8-lane slot of a
// wave takes two rows, base + s and base + s + 8, and issues the loads of
// both rows' first 32 entries before the first gather -- twice the rows per
// wave of k_csr_matvec<8>, in half the waves.  Entries beyond the 32nd (a
// general CSR may have such rows) are added in the same per-lane order.
// The sums are bit-identical to k_csr_matvec<8, false>.  Measured on the
// 24-qubit Heisenberg CSR (25 entries per row): 2.61 -> 2.47 ms; four rows
// per slot: 2.65 ms.
__global__ __launch_bounds__(256) void k_csr_matvec_short(uint64_t nrows, uint64_t ncols, uint64_t nnz,
                                                          uint32_t nblocks_xcd,
                                                          const uint64_t *__restrict__ indptr,
                                                          const uint64_t *__restrict__ indices,
                                                          const double2 *__restrict__ data,
                                                          const double2 *__restrict__ x,
                                                          double2 *__restrict__ y) {
    constexpr int LPR = 8, U = 4, H = 2;
    uint32_t b = blockIdx.x;
    if (nblocks_xcd) b = (b % 8) * nblocks_xcd + b / 8;
    const uint32_t wave = threadIdx.x / 64, slot = (threadIdx.x % 64) / LPR;
    const uint32_t lane = threadIdx.x % LPR;
    const uint64_t base = (uint64_t(b) * 4 + wave) * (LPR * H);
    uint64_t c[H][U], k0[H], k1[H];
    double2 a[H][U], acc[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        acc[h] = make_double2(0.0, 0.0);
        const uint64_t row = base + slot + LPR * h;
        k0[h] = k1[h] = 0;
        if (row < nrows) {
            k0[h] = min(indptr[row], nnz);
            k1[h] = min(indptr[row + 1], nnz);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t k = k0[h] + lane + u * LPR;
            c[h][u] = ~uint64_t(0);
            a[h][u] = make_double2(0.0, 0.0);
            if (k < k1[h]) {
                c[h][u] = indices[k];
                const dvec2 av = *reinterpret_cast<const dvec2 *>(data + k);
                a[h][u] = make_double2(av.x, av.y);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (c[h][u] < ncols) {
                const double2 bx = x[c[h][u]];
                acc[h].x = fma(a[h][u].x, bx.x, acc[h].x);
                acc[h].x = fma(-a[h][u].y, bx.y, acc[h].x);
                acc[h].y = fma(a[h][u].x, bx.y, acc[h].y);
                acc[h].y = fma(a[h][u].y, bx.x, acc[h].y);
            }
        }
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const uint64_t row = base + slot + LPR * h;
        if (k1[h] > k0[h] + U * LPR) {
            // a row longer than U * LPR entries: the rest, in the same order
            for (uint64_t k = k0[h] + lane + U * LPR; k < k1[h]; k += LPR) {
                const uint64_t cc = indices[k];
                if (cc < ncols) {
                    const double2 av = data[k], bx = x[cc];
                    acc[h].x = fma(av.x, bx.x, acc[h].x);
                    acc[h].x = fma(-av.y, bx.y, acc[h].x);
                    acc[h].y = fma(av.x, bx.y, acc[h].y);
                    acc[h].y = fma(av.y, bx.x, acc[h].y);
                }
            }
        }
#pragma unroll
        for (int off = LPR / 2; off > 0; off /= 2) {
            acc[h].x += __shfl_xor(acc[h].x, off, LPR);
            acc[h].y += __shfl_xor(acc[h].y, off, LPR);
        }
        if (row < nrows && lane == 0) y[row] = acc[h];
    }
}

// k_csr_matvec_staged (the walker of qr_mv.h over a general CSR): which band
// of x the resident waves read.  In a column-sorted row, chunk c
// covers about the same band of x in every row (1/31 of x, 0.5 MB, for the
// 2000-term CSR), so waves working on the same chunk index read the same
// band, and the XCD's 4 MB L2 serves them.  Blocks start whenever a slot
// frees, and in steady state the resident ones are spread evenly over their
// rows: block i (in launch order on its XCD, of K resident per XCD) starts
// about i/K of a block's lifetime after block 0.  So block i starts its rows
// at chunk i*n/K (mod n, for n chunks) and wraps around: a block then sits at
// the same chunk index as the others when it runs.  Measured on the
// 2000-term CSR: 10.20 ms without the offset, 9.68 ms with it (K = 96, the
// residency on MI355X; 48: 10.99 ms, 192: 10.23 ms), 8.97 ms with every
// gather served from L2 (a timing-only build); a persistent grid (one
// resident set of blocks looping over the rows, all starting at chunk 0)
// took 10.29 ms, its waves drifting apart.  The summation order depends only
// on the block index, so results are reproducible.
// One block per 32 rows; block index remapped so that each XCD works on one
// contiguous range of rows (blocks go to the 8 XCDs round-robin).
template <int C>
__global__ __launch_bounds__(256) void k_csr_matvec_staged(uint64_t nrows, uint64_t ncols, uint64_t nnz,
                                                           uint32_t nblocks_xcd, uint32_t k_res,
                                                           const uint64_t *__restrict__ indptr,
                                                           const uint64_t *__restrict__ indices,
                                                           const double2 *__restrict__ data,
                                                           const double2 *__restrict__ x,
                                                           double2 *__restrict__ y) {
    using S = MvStagedShape<C>;
    extern __shared__ __align__(16) unsigned char mv_lds[];
    uint32_t b = blockIdx.x;
    if (nblocks_xcd) b = (b % 8) * nblocks_xcd + b / 8;
    const int wave = int(threadIdx.x / 64);
    uint64_t *s_idx = reinterpret_cast<uint64_t *>(mv_lds + size_t(wave) * S::kWaveBytes);
    dvec2 *s_dat = reinterpret_cast<dvec2 *>(s_idx + S::kRows * S::kIdxStride);
    // row bounds clamped to [0, nnz), so that a malformed indptr cannot read out
    // of bounds
    auto bounds = [indptr, nnz](uint64_t r, uint64_t *k0, uint64_t *len) {
        const uint64_t a = min(indptr[r], nnz), e = min(indptr[r + 1], nnz);
        *k0 = a;
        *len = e > a ? e - a : 0;
    };
    mv_staged_rows<C>((uint64_t(b) * 4 + uint64_t(wave)) * S::kRows, nrows, ncols, bounds, indices,
                      data, x, y, s_idx, s_dat, blockIdx.x / 8, k_res);
}

// Blocks of k_csr_matvec_staged resident at once on the current device (per
// device, computed once).
static int mv_staged_resident(int *out) {
    static std::mutex mu;
    static int cache[kMaxDevices];
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!cache[dev]) {
        int per_cu = 0, cus = 0;
        QR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(k_csr_matvec_staged<kMvStagedChunk>), 256,
            4 * MvStagedShape<kMvStagedChunk>::kWaveBytes));
        QR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cache[dev] = std::max(1, per_cu) * std::max(1, cus);
    }
    *out = cache[dev];
    return QR_OK;
}

int launch_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                      const uint64_t *d_indices, const double *d_data, const double *d_x,
                      double *d_y, hipStream_t s) {
    const uint64_t avg = nrows ? nnz / nrows : 0;
    if (avg >= kMvStagedMinRow) {
        using Sh = MvStagedShape<kMvStagedChunk>;
        int resident = 0;
        if (int rc = mv_staged_resident(&resident)) return rc;
        const uint32_t k_res = uint32_t(std::max(1, resident / 8));  // per XCD
        const uint64_t nb = ((nrows + Sh::kRows - 1) / Sh::kRows + 3) / 4;
        if (nb > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR matvec grid too large");
        const uint32_t nbx = (nb % 8 == 0) ? uint32_t(nb / 8) : 0;
        hipLaunchKernelGGL(k_csr_matvec_staged<kMvStagedChunk>, dim3(unsigned(nb)), dim3(256),
                           4 * Sh::kWaveBytes, s, nrows, ncols, nnz, nbx, k_res, d_indptr, d_indices,
                           reinterpret_cast<const double2 *>(d_data),
                           reinterpret_cast<const double2 *>(d_x), reinterpret_cast<double2 *>(d_y));
        QR_HIP_CHECK(hipGetLastError());
        return QR_OK;
    }
    // ~4+ entries per lane, at most 8 lanes per row (8 rows per wave; 4 and 16
    // to 32 lanes per row measured slower at G = 2000, and 16 no faster at
    // G = 25)
    int lpr = 1;
    while (lpr < 8 && uint64_t(lpr) * 4 <= avg) lpr *= 2;
    const uint64_t threads = nrows * uint64_t(lpr);
    const uint64_t nblocks = (threads + 255) / 256;
    if (nblocks > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR matvec grid too large");
    const uint32_t nbx = (nblocks % 8 == 0) ? uint32_t(nblocks / 8) : 0;
    auto x = reinterpret_cast<const double2 *>(d_x);
    auto data = reinterpret_cast<const double2 *>(d_data);
    auto y = reinterpret_cast<double2 *>(d_y);
    const dim3 grid(static_cast<unsigned>(nblocks)), block(256);
    const bool nt = avg >= kMvNontemporalMinRow;
    if (lpr == 8 && avg <= 32) {
        // 16..32 entries per row on average: two rows per 8-lane slot
        const uint64_t nb = (nrows + 63) / 64;
        if (nb > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "CSR matvec grid too large");
        const uint32_t nbx_short = (nb % 8 == 0) ? uint32_t(nb / 8) : 0;
        hipLaunchKernelGGL(k_csr_matvec_short, dim3(unsigned(nb)), dim3(256), 0, s, nrows, ncols,
                           nnz, nbx_short, d_indptr, d_indices,
                           reinterpret_cast<const double2 *>(d_data),
                           reinterpret_cast<const double2 *>(d_x), reinterpret_cast<double2 *>(d_y));
        QR_HIP_CHECK(hipGetLastError());
        return QR_OK;
    }
#define QR_MV_CASE(L)                                                                        \
    case L:                                                                                  \
        if (nt)                                                                              \
            hipLaunchKernelGGL((k_csr_matvec<L, true>), grid, block, 0, s, nrows, ncols, nnz, \
                               nbx, d_indptr, d_indices, data, x, y);                        \
        else                                                                                 \
            hipLaunchKernelGGL((k_csr_matvec<L, false>), grid, block, 0, s, nrows, ncols, nnz, \
                               nbx, d_indptr, d_indices, data, x, y);                        \
        break;
    switch (lpr) {
        QR_MV_CASE(1)
        QR_MV_CASE(2)
        QR_MV_CASE(4)
        default: QR_MV_CASE(8)
    }
#undef QR_MV_CASE
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

}  // namespace qr
