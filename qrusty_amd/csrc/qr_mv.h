// qr_mv.h -- the row-staged CSR x vector walker shared by k_csr_matvec_staged
// (qr_csr.hip) and the batch's matvec pass (qr_batch.hip), for rows of at
// least kMvStagedMinRow entries (accel::rowwise::spmat_dot_densevec,
// qrusty/src/accel.rs:338-370).
//
// Long rows: the row stream and the x gathers want different lane layouts.  Streams run fastest when every
// load instruction reads one contiguous run (measured on the 20-qubit,
// 2000-term CSR with the gathers left out: 7.77 ms with a whole wave per row,
// 9.23 ms with 8 lanes per row, where an instruction reads 8 rows in 64/128 B
// pieces, half of them straddling two lines).  Gathers want the 8 lanes that
// read one x line in the same instruction: for the Pauli CSR, rows that differ
// in their low 3 bits read the same line at the same position of their
// (column-sorted) rows, so a gather instruction should cover 8 rows at a few
// positions each -- with a wave per row each gather instruction touches 64
// lines and the matvec takes 16.2 ms.
//
// So each wave owns 8 consecutive rows and walks them in chunks of C
// positions: it streams each row's chunk with contiguous loads (one row per
// instruction for C = 64), writes it to its own padded LDS region, reads it
// back transposed -- lane (s, t) takes positions p(s, u) of row t -- and
// gathers x there.  The next chunk's stream loads are issued after the
// gathers, so that waiting on a gather does not wait on them.  The paddings
// keep both LDS access patterns free of bank conflicts.  The 4 waves of a
// block walk their chunks in step (a barrier per chunk, over the block's
// longest rows), so that they read the same band of x (see qr_csr.hip).
//
// The walker reads memory only through pointers cast to the global address
// space: callers that take their arrays from descriptors in memory would
// otherwise access them with flat instructions, which count against both
// vmcnt and lgkmcnt.
#pragma once

#include "qr_internal.h"

namespace qr {

// FP64 pair as a vector type: 16-byte loads and stores
typedef double dvec2 __attribute__((ext_vector_type(2)));

constexpr uint64_t kMvStagedMinRow = 256;
constexpr int kMvStagedChunk = 64;  // positions per row and chunk (32 measured 10% slower)

template <int C>
struct MvStagedShape {
    static constexpr int kRows = 8;              // rows per wave
    static constexpr int kRowsPerLoad = 64 / C;  // rows covered by one load instruction
    static constexpr int kLoads = kRows / kRowsPerLoad;
    static constexpr int kU = C / 8;             // positions per lane per chunk (gather layout)
    // C = 64: lane (s, t) takes positions 8s..8s+7; C = 32: positions s, s+8, s+16, s+24
    static constexpr bool kRun = C == 64;
    static constexpr int kIdxStride = C == 64 ? C + 1 : C + 2;   // u64 slots per row
    static constexpr int kDatStride = C + 1;                     // double2 slots per row
    static constexpr size_t kWaveBytes = size_t(kRows) * (kIdxStride * 8 + kDatStride * 16);
    __device__ static constexpr int pos(int s, int u) { return kRun ? s * kU + u : u * 8 + s; }
};

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
    return (uint64_t(hi) << 32) | lo;
}

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *as_global(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// One wave's 8 rows r0..r0+7 (rows >= nrows are empty).  bounds(row, &k0,
// &len) gives a row's first entry and length, within the arrays.  The first
// chunk is i_blk * n / k_res (mod n, for n chunks); k_res = 0: chunk 0.
template <int C, typename Bounds>
__device__ __forceinline__ void mv_staged_rows(uint64_t r0, uint64_t nrows, uint64_t ncols, Bounds bounds,
                                               const uint64_t *indices_p, const double2 *data_p,
                                               const double2 *x_p, double2 *y_p, uint64_t *s_idx,
                                               dvec2 *s_dat, uint64_t i_blk, uint64_t k_res) {
    const auto indices = as_global(indices_p);
    const auto data = as_global(data_p);
    // x and y as FP64 pairs of builtin vector type (HIP's double2 class does not
    // copy across address spaces)
    const auto x = as_global(reinterpret_cast<const dvec2 *>(x_p));
    const auto y = as_global(reinterpret_cast<dvec2 *>(y_p));
    using S = MvStagedShape<C>;
    constexpr int R = S::kRows;
    const int l = int(threadIdx.x % 64);
    const int s = l / 8, t = l % 8;

    // row bounds: lane j < 8 holds row r0 + j
    uint64_t k0l = 0, lenl = 0;
    if (l < R && r0 + l < nrows) bounds(r0 + l, &k0l, &lenl);
    uint64_t k0[R], len[R], maxlen = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        k0[j] = readlane_u64(k0l, j);
        len[j] = readlane_u64(lenl, j);
        maxlen = max(maxlen, len[j]);
    }
    const uint64_t lent = __shfl(lenl, t);  // length of this lane's gather row
    // the block walks the chunk count of its longest rows, every wave in step
    // with the others (a barrier per chunk); chunks past a wave's rows load
    // and add nothing.  Every wave of the block calls this function for the
    // same number of row groups.
    uint64_t nchunks = (maxlen + C - 1) / C;
    {
        __shared__ unsigned long long nch_s[4];
        if (l == 0) nch_s[threadIdx.x / 64] = nchunks;
        __syncthreads();
        nchunks = max(max(nch_s[0], nch_s[1]), max(nch_s[2], nch_s[3]));
        __syncthreads();
    }

    // stream registers: load instruction q covers row q * kRowsPerLoad + l / C
    uint64_t si[S::kLoads];
    dvec2 sd[S::kLoads];
    auto stream = [&](uint64_t c) {
#pragma unroll
        for (int q = 0; q < S::kLoads; ++q) {
            const uint64_t kb = S::kRowsPerLoad == 1 ? k0[q] : (l < C ? k0[2 * q] : k0[2 * q + 1]);
            const uint64_t lj = S::kRowsPerLoad == 1 ? len[q] : (l < C ? len[2 * q] : len[2 * q + 1]);
            const uint64_t p = c * C + uint64_t(l % C);
            if (p < lj) {
                si[q] = __builtin_nontemporal_load(indices + kb + p);
                sd[q] = __builtin_nontemporal_load(
                    reinterpret_cast<const __attribute__((address_space(1))) dvec2 *>(data + kb + p));
            }
        }
    };
#pragma unroll
    for (int q = 0; q < S::kLoads; ++q) {
        si[q] = ~uint64_t(0);
        sd[q] = dvec2{0.0, 0.0};
    }
    // first chunk: this block's share of the resident blocks' stagger
    const uint64_t c0 = nchunks && k_res ? (i_blk % k_res) * nchunks / k_res : 0;
    if (nchunks) stream(c0);

    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t cc = 0, c = c0; cc < nchunks; ++cc) {
#pragma unroll
        for (int q = 0; q < S::kLoads; ++q) {
            const int j = q * S::kRowsPerLoad + (S::kRowsPerLoad == 1 ? 0 : l / C);
            s_idx[j * S::kIdxStride + l % C] = si[q];
            s_dat[j * S::kDatStride + l % C] = sd[q];
        }
        // LDS writes of other lanes visible to this wave's reads below
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t gi[S::kU];
        dvec2 gd[S::kU];
        double2 gx[S::kU];
        bool ok[S::kU];
#pragma unroll
        for (int u = 0; u < S::kU; ++u) {
            const int p = S::pos(s, u);
            gi[u] = s_idx[t * S::kIdxStride + p];
            gd[u] = s_dat[t * S::kDatStride + p];
            ok[u] = (c * C + uint64_t(p) < lent) && gi[u] < ncols;
        }
#pragma unroll
        for (int u = 0; u < S::kU; ++u) {
            gx[u] = make_double2(0.0, 0.0);
            if (ok[u]) {
                const dvec2 v = x[gi[u]];
                gx[u] = make_double2(v.x, v.y);
            }
        }
        // the next chunk's stream loads go out behind the gathers
        __builtin_amdgcn_sched_barrier(0);
        const uint64_t cn = c + 1 == nchunks ? 0 : c + 1;
        if (cc + 1 < nchunks) stream(cn);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < S::kU; ++u) {
            if (ok[u]) {
                acc.x = fma(gd[u].x, gx[u].x, acc.x);
                acc.x = fma(-gd[u].y, gx[u].y, acc.x);
                acc.y = fma(gd[u].x, gx[u].y, acc.y);
                acc.y = fma(gd[u].y, gx[u].x, acc.y);
            }
        }
        // this chunk's LDS reads done before the next chunk's writes, and the
        // block's waves in step: measured on the 20-qubit, 2000-term CSR,
        // 9.19-9.41 ms against 9.64-9.67 with each wave walking alone
        __syncthreads();
        c = cn;
    }
#pragma unroll
    for (int off = 8; off < 64; off *= 2) {
        acc.x += __shfl_xor(acc.x, off);
        acc.y += __shfl_xor(acc.y, off);
    }
    if (s == 0 && r0 + t < nrows) y[r0 + t] = dvec2{acc.x, acc.y};
}

}  // namespace qr
