// qr_spadd.hip -- a*A + b*B of two CSR matrices of the same shape on the
// device: pyqrusty's a_spmat_p_b_spmat (pyqrusty/src/lib.rs:499-517), i.e.
// sprs CsMat::scale on clones and `&am1 + &am2` (sprs csmat_binop).
//
// Each row is the merge of the two rows' ascending columns: an entry in both
// gets a*v1 + b*v2, one in A only a*v1 + 0, one in B only 0 + b*v2 (complex
// products as num-complex, no fused multiply-add), and results equal to zero
// are dropped.  Two passes over the rows -- count, then (after a scan of the
// counts into indptr) fill -- with one thread per row.
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

__global__ __launch_bounds__(256) void k_spadd_count(uint64_t nrows, SpAddIn in,
                                                     uint64_t *__restrict__ counts) {
    const uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= nrows) return;
    uint64_t n = 0;
    merge_row(r, in, [&](uint64_t, double2) { ++n; });
    counts[r] = n;
}

__global__ __launch_bounds__(256) void k_spadd_fill(uint64_t nrows, SpAddIn in,
                                                    const uint64_t *__restrict__ ipo,
                                                    uint64_t *__restrict__ ixo,
                                                    double2 *__restrict__ dout) {
    const uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= nrows) return;
    uint64_t k = ipo[r];
    merge_row(r, in, [&](uint64_t c, double2 v) {
        ixo[k] = c;
        dout[k] = v;
        ++k;
    });
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
    const unsigned grid = unsigned((nrows + 255) / 256);
    size_t tmp_bytes = 0;
    QR_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, ipo, ipo + 1, nrows, s));
    // per-row counts, then their inclusive scan into indptr[1..nrows]
    void *tmp = nullptr;
    const size_t cbytes = (size_t(nrows) * sizeof(uint64_t) + 255) & ~size_t(255);
    QR_HIP_CHECK(hipMallocAsync(&tmp, cbytes + tmp_bytes, s));
    uint64_t *counts = static_cast<uint64_t *>(tmp);
    hipLaunchKernelGGL(k_set_zero_u64, dim3(1), dim3(1), 0, s, ipo);
    hipLaunchKernelGGL(k_spadd_count, dim3(grid), dim3(256), 0, s, nrows, in, counts);
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
    const unsigned grid = unsigned((nrows + 255) / 256);
    hipLaunchKernelGGL(k_spadd_fill, dim3(grid), dim3(256), 0, s, nrows, in, ipo, ixo,
                       reinterpret_cast<double2 *>(dout));
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

}  // namespace qr
