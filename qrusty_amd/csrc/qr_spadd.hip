// qr_spadd.hip -- a*A + b*B of two CSR matrices of the same shape on the
// device: pyqrusty's a_spmat_p_b_spmat (pyqrusty/src/lib.rs:499-517), i.e.
// sprs CsMat::scale on clones and `&am1 + &am2` (sprs csmat_binop).
//
// Each row is the merge of the two rows' ascending columns: an entry in both
// gets a*v1 + b*v2, one in A only a*v1 + 0, one in B only 0 + b*v2 (complex
// products as num-complex, no fused multiply-add), and results equal to zero
// are dropped.  Two passes over the rows -- count, then (after a scan of the
// counts into indptr) fill.
//
// A block takes a tile of consecutive rows.  Their entries are contiguous in
// both inputs, so the block loads them coalesced into LDS (column and the
// already scaled value a*v1 or b*v2), a group of lanes per row merges in LDS
// (each lane from its own merge-path split of the row), and
// the fill pass writes the tile's merged entries to LDS and from there
// coalesced to the output, whose tile range is contiguous too.  A tile whose
// entries exceed the LDS capacity (long or uneven rows) is merged from global
// memory, one thread per row.
#include "qr_internal.h"

#include <hipcub/device/device_scan.hpp>

namespace qr {

__device__ __forceinline__ double2 cmul2(double2 a, double2 v) {
    return make_double2(a.x * v.x - a.y * v.y, a.x * v.y + a.y * v.x);
}

// Walks the merge of row r; for every kept entry calls emit(column, value).
// Row bounds are clamped to each matrix's nnz, so a malformed CSR cannot make
// the merge read out of bounds.
struct SpAddIn {
    double2 a, b;
    uint64_t nnz1, nnz2;
    const uint64_t *ip1, *ix1, *ip2, *ix2;
    const double2 *d1, *d2;
};

template <typename Emit>
__device__ __forceinline__ void merge_row(uint64_t r, const SpAddIn &in, Emit emit) {
    const double2 a = in.a, b = in.b;
    const uint64_t *__restrict__ ix1 = in.ix1;
    const uint64_t *__restrict__ ix2 = in.ix2;
    const double2 *__restrict__ d1 = in.d1;
    const double2 *__restrict__ d2 = in.d2;
    const uint64_t i1 = in.ip1[r + 1] < in.nnz1 ? in.ip1[r + 1] : in.nnz1;
    const uint64_t j1 = in.ip2[r + 1] < in.nnz2 ? in.ip2[r + 1] : in.nnz2;
    uint64_t i = in.ip1[r] < i1 ? in.ip1[r] : i1, j = in.ip2[r] < j1 ? in.ip2[r] : j1;
    const double2 zero = make_double2(0.0, 0.0);
    while (i < i1 || j < j1) {
        uint64_t col;
        double2 v;
        if (j >= j1 || (i < i1 && ix1[i] < ix2[j])) {
            col = ix1[i];
            const double2 p = cmul2(a, d1[i]);
            v = make_double2(p.x + zero.x, p.y + zero.y);
            ++i;
        } else if (i >= i1 || ix2[j] < ix1[i]) {
            col = ix2[j];
            const double2 q = cmul2(b, d2[j]);
            v = make_double2(zero.x + q.x, zero.y + q.y);
            ++j;
        } else {
            col = ix1[i];
            const double2 p = cmul2(a, d1[i]), q = cmul2(b, d2[j]);
            v = make_double2(p.x + q.x, p.y + q.y);
            ++i;
            ++j;
        }
        if (v.x != 0.0 || v.y != 0.0) emit(col, v);
    }
}

#ifndef QR_SPADD_CAP
#define QR_SPADD_CAP 512
#endif
// Threads per block: measured on the 24-qubit H - e*I (count + fill, ms) with
// 512-entry tiles, 256 threads 3.2 + 5.4, 128 threads 2.7 + 6.3.  The two
// passes need not share a tiling (for a CSR whose indptr ascends within nnz,
// each route of the merge gives a row the same entries), so each takes its
// faster one.  Count pass, threads / tile entries: 128 / 512 2.52-2.64 ms,
// 64 / 256 2.60, 64 / 512 2.85, 256 / 1024 2.98, 128 / 256 3.10,
// 128 / 1024 3.45.
#ifndef QR_SPADD_COUNT_THREADS
#define QR_SPADD_COUNT_THREADS 128
#endif
#ifndef QR_SPADD_COUNT_CAP
#define QR_SPADD_COUNT_CAP 512
#endif
constexpr int kSpCountThreads = QR_SPADD_COUNT_THREADS, kSpFillThreads = 256;
constexpr int kSpCountCap = QR_SPADD_COUNT_CAP;
constexpr int kSpCap = QR_SPADD_CAP;  // staged input entries (A + B) per tile
static_assert(kSpCap % kSpFillThreads == 0 && kSpCountCap % kSpCountThreads == 0,
              "whole staging rounds per thread");

__device__ __forceinline__ uint64_t clampu(uint64_t v, uint64_t lo, uint64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Tile of rows [r0, r0 + rows_per_tile).  kFill = false writes counts[r];
// kFill = true writes the merged rows at ipo[r].
template <bool kFill, int kT, int kCap>
__global__ __launch_bounds__(kT) void k_spadd_tile(uint64_t nrows, uint32_t rows_per_tile,
                                                           uint32_t lpr,
                                                           SpAddIn in, uint64_t *__restrict__ counts,
                                                           const uint64_t *__restrict__ ipo,
                                                           uint64_t *__restrict__ ixo,
                                                           double2 *__restrict__ dout) {
    __shared__ uint64_t sc[kCap];
    __shared__ double2 sv[kCap];
    __shared__ uint64_t oc[kFill ? kCap : 1];
    __shared__ double2 ov[kFill ? kCap : 1];
    const uint32_t tid = threadIdx.x;
    const uint64_t r0 = uint64_t(blockIdx.x) * rows_per_tile;
    if (r0 >= nrows) return;
    const uint64_t r1 = r0 + rows_per_tile < nrows ? r0 + rows_per_tile : nrows;
    const uint64_t a0 = clampu(in.ip1[r0], 0, in.nnz1), a1 = clampu(in.ip1[r1], a0, in.nnz1);
    const uint64_t b0 = clampu(in.ip2[r0], 0, in.nnz2), b1 = clampu(in.ip2[r1], b0, in.nnz2);
    const uint64_t ea = a1 - a0, eb = b1 - b0;
    if (ea + eb > uint64_t(kCap)) {
        // global-memory merge, one thread per row.  For a CSR whose indptr
        // ascends within nnz it gives each row what the LDS merge gives; the
        // two passes tile the rows differently, so for a malformed one a row's
        // count may come from the other route, and the fill stops at the
        // row's end either way.
        for (uint64_t r = r0 + tid; r < r1; r += kT) {
            if (kFill) {
                uint64_t k = ipo[r];
                const uint64_t kend = ipo[r + 1];
                merge_row(r, in, [&](uint64_t c, double2 v) {
                    if (k < kend) {
                        ixo[k] = c;
                        dout[k] = v;
                    }
                    ++k;
                });
            } else {
                uint64_t n = 0;
                merge_row(r, in, [&](uint64_t, double2) { ++n; });
                counts[r] = n;
            }
        }
        return;
    }
    {
        // every load of this thread in flight at once: slot k < ea is A's
        // entry a0 + k, slot ea <= k < ea + eb B's entry b0 + k - ea
        constexpr int kPer = kCap / kT;
        uint64_t c[kPer];
        double2 v[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const uint32_t k = tid + u * kT;
            if (k < ea) {
                c[u] = __builtin_nontemporal_load(in.ix1 + a0 + k);
                v[u].x = __builtin_nontemporal_load(&in.d1[a0 + k].x);
                v[u].y = __builtin_nontemporal_load(&in.d1[a0 + k].y);
            } else if (k < ea + eb) {
                c[u] = __builtin_nontemporal_load(in.ix2 + b0 + (k - ea));
                v[u].x = __builtin_nontemporal_load(&in.d2[b0 + (k - ea)].x);
                v[u].y = __builtin_nontemporal_load(&in.d2[b0 + (k - ea)].y);
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const uint32_t k = tid + u * kT;
            if (k < ea + eb) {
                sc[k] = c[u];
                sv[k] = cmul2(k < ea ? in.a : in.b, v[u]);
            }
        }
    }
    __syncthreads();
    const uint64_t o0 = kFill ? ipo[r0] : 0;
    // lpr lanes per row (a power of two <= 64, so a row's lanes share a wave)
    const uint32_t lane = tid & (lpr - 1);
    for (uint64_t r = r0 + tid / lpr; r < r1; r += kT / lpr) {
        // row bounds inside the staged ranges: B's entries sit after A's
        const uint32_t ia = uint32_t(clampu(in.ip1[r], a0, a1) - a0);
        const uint32_t ia1 = uint32_t(clampu(in.ip1[r + 1], a0 + ia, a1) - a0);
        const uint32_t jb = uint32_t(ea + clampu(in.ip2[r], b0, b1) - b0);
        const uint32_t jb1 = uint32_t(ea + clampu(in.ip2[r + 1], b0 + (jb - ea), b1) - b0);
        const uint32_t na = ia1 - ia, nb = jb1 - jb, m = na + nb;
        const uint32_t seg = (m + lpr - 1) / lpr;
        // merge-path split (A first on equal columns) at the diagonals of this
        // lane and of the next
        auto split = [&](uint32_t d) {
            uint32_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sc[ia + mid] <= sc[jb + d - 1 - mid]) lo = mid + 1;
                else hi = mid;
            }
            return lo;
        };
        const uint32_t d0 = lane * seg < m ? lane * seg : m;
        const uint32_t dE = d0 + seg < m ? d0 + seg : m;
        const uint32_t s0 = split(d0), sE = split(dE);
        uint32_t i = ia + s0, j = jb + (d0 - s0);
        const uint32_t iE = ia + sE, jE = jb + (dE - sE);
        // a B entry paired with the last A entry before this lane's start was
        // taken by the lane that holds that A entry
        if (i > ia && j < jb1 && sc[i - 1] == sc[j]) ++j;
        auto walk = [&](auto emit) {
            uint32_t wi = i, wj = j;
            while (wi < iE || wj < jE) {
                bool take_a, take_b;
                uint64_t col;
                if (wi >= iE) {
                    take_a = false;
                    take_b = true;
                    col = sc[wj];
                } else {
                    const uint64_t ca = sc[wi];
                    // the B entry at this lane's end may pair with its last A entry
                    const bool b_ok = wj < jE || (wj < jb1 && wi + 1 == iE);
                    const uint64_t cb = b_ok ? sc[wj] : ca;
                    take_a = !b_ok || ca <= cb;
                    take_b = b_ok && cb <= ca;
                    col = take_a ? ca : cb;
                }
                double2 v;
                // one side only: + 0 as the reference's binop does (a -0.0
                // component becomes +0.0)
                if (take_a && take_b) {
                    const double2 p = sv[wi++], q = sv[wj++];
                    v = make_double2(p.x + q.x, p.y + q.y);
                } else if (take_a) {
                    const double2 p = sv[wi++];
                    v = make_double2(p.x + 0.0, p.y + 0.0);
                } else {
                    const double2 q = sv[wj++];
                    v = make_double2(0.0 + q.x, 0.0 + q.y);
                }
                if (v.x != 0.0 || v.y != 0.0) emit(col, v);
            }
        };
        // kept entries of this lane; the fill pass walks again to emit them
        // from the slot after the lanes before it
        uint32_t n = 0;
        walk([&](uint64_t, double2) { ++n; });
        // inclusive sum of n over the row's lanes
        uint32_t incl = n;
        for (uint32_t o = 1; o < lpr; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, lpr);
            if (lane >= o) incl += t;
        }
        if (!kFill) {
            if (lane == lpr - 1) counts[r] = incl;
        } else {
            uint64_t k = ipo[r] - o0 + (incl - n);
            walk([&](uint64_t col, double2 v) {
                if (k < uint64_t(kCap)) {
                    oc[k] = col;
                    ov[k] = v;
                }
                ++k;
            });
        }
    }
    if (kFill) {
        __syncthreads();
        uint64_t total = ipo[r1] - o0;
        total = total < uint64_t(kCap) ? total : uint64_t(kCap);
        for (uint32_t k = tid; k < total; k += kT) {
            __builtin_nontemporal_store(oc[k], ixo + o0 + k);
            __builtin_nontemporal_store(ov[k].x, &dout[o0 + k].x);
            __builtin_nontemporal_store(ov[k].y, &dout[o0 + k].y);
        }
    }
}

// Rows per tile: the staged capacity over the mean merged row length, with
// room for rows up to 1/0.85 times the mean.  Lanes per row: enough that the
// tile's rows keep the block's threads busy (a power of two, at most 64).
struct SpTiling {
    uint32_t rows, lpr;
};
static SpTiling spadd_tiling(uint64_t nrows, uint64_t nnz1, uint64_t nnz2, int threads, int cap) {
    const double mean = double(nnz1 + nnz2) / double(nrows ? nrows : 1);
    const double r = mean > 0 ? 0.85 * cap / mean : double(threads);
    const uint32_t rows = r >= threads ? uint32_t(threads) : (r < 1 ? 1u : uint32_t(r));
    uint32_t lpr = 1;
    while (lpr < 64 && lpr * rows < uint32_t(threads)) lpr <<= 1;
    return SpTiling{rows, lpr};
}

__global__ void k_set_zero_u64(uint64_t *p) { *p = 0; }

static SpAddIn spadd_in(const double *a, uint64_t nnz1, const uint64_t *ip1, const uint64_t *ix1,
                        const double *d1, const double *b, uint64_t nnz2, const uint64_t *ip2,
                        const uint64_t *ix2, const double *d2) {
    return SpAddIn{make_double2(a[0], a[1]), make_double2(b[0], b[1]), nnz1, nnz2, ip1, ix1, ip2,
                   ix2, reinterpret_cast<const double2 *>(d1), reinterpret_cast<const double2 *>(d2)};
}

int launch_spadd_count(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *ip1,
                       const uint64_t *ix1, const double *d1, const double *b, uint64_t nnz2,
                       const uint64_t *ip2, const uint64_t *ix2, const double *d2, uint64_t *ipo,
                       uint64_t *nnz_out, hipStream_t s) {
    const SpAddIn in = spadd_in(a, nnz1, ip1, ix1, d1, b, nnz2, ip2, ix2, d2);
    size_t tmp_bytes = 0;
    QR_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, ipo, ipo + 1, nrows, s));
    // per-row counts, then their inclusive scan into indptr[1..nrows]
    void *tmp = nullptr;
    const size_t cbytes = (size_t(nrows) * sizeof(uint64_t) + 255) & ~size_t(255);
    QR_HIP_CHECK(hipMallocAsync(&tmp, cbytes + tmp_bytes, s));
    uint64_t *counts = static_cast<uint64_t *>(tmp);
    const SpTiling tl = spadd_tiling(nrows, nnz1, nnz2, kSpCountThreads, kSpCountCap);
    const uint64_t ntiles = (nrows + tl.rows - 1) / tl.rows;
    if (ntiles > 0x7FFFFFFFull) {
        (void)hipFreeAsync(tmp, s);
        return fail(QR_ERR_SIZE, "sparse add: too many row tiles");
    }
    hipLaunchKernelGGL(k_set_zero_u64, dim3(1), dim3(1), 0, s, ipo);
    hipLaunchKernelGGL((k_spadd_tile<false, kSpCountThreads, kSpCountCap>), dim3(unsigned(ntiles)), dim3(kSpCountThreads), 0, s, nrows,
                       tl.rows, tl.lpr, in, counts, nullptr, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipcub::DeviceScan::InclusiveSum(static_cast<unsigned char *>(tmp) + cbytes, tmp_bytes,
                                             counts, ipo + 1, nrows, s);
    (void)hipFreeAsync(tmp, s);
    if (e != hipSuccess) return hip_fail(e, "sparse add count / scan");
    QR_HIP_CHECK(hipMemcpyAsync(nnz_out, ipo + nrows, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    QR_HIP_CHECK(hipStreamSynchronize(s));
    return QR_OK;
}

int launch_spadd_fill(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *ip1,
                      const uint64_t *ix1, const double *d1, const double *b, uint64_t nnz2,
                      const uint64_t *ip2, const uint64_t *ix2, const double *d2,
                      const uint64_t *ipo, uint64_t *ixo, double *dout, hipStream_t s) {
    const SpAddIn in = spadd_in(a, nnz1, ip1, ix1, d1, b, nnz2, ip2, ix2, d2);
    const SpTiling tl = spadd_tiling(nrows, nnz1, nnz2, kSpFillThreads, kSpCap);
    const uint64_t ntiles = (nrows + tl.rows - 1) / tl.rows;
    if (ntiles > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "sparse add: too many row tiles");
    hipLaunchKernelGGL((k_spadd_tile<true, kSpFillThreads, kSpCap>), dim3(unsigned(ntiles)), dim3(kSpFillThreads), 0, s, nrows,
                       tl.rows, tl.lpr, in, nullptr, ipo, ixo, reinterpret_cast<double2 *>(dout));
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

}  // namespace qr
