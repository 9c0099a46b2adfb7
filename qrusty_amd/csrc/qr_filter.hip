// qr_filter.hip -- tolerance filters of a device CSR:
//   count_zeros(tol):     entries with |v| <= tol   (qrusty/src/util.rs:144-152)
//   eliminate_zeros(tol): the CSR of the entries with |v| > tol
//                         (qrusty/src/util.rs:154-171)
// |v| is num-complex's norm(), hypot(re, im).  Both are streaming passes:
// count is a reduction over a flag iterator; eliminate is an exclusive scan of
// the keep flags (one position per entry, plus the total) followed by a
// scatter of the kept entries and indptr_out[r] = pos[indptr[r]].
#include "qr_internal.h"

#include <hipcub/device/device_reduce.hpp>
#include <hipcub/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace qr {

struct KeepFlag {
    const double2 *d;
    uint64_t nnz;
    double tol;
    bool keep_small;  // true: flag |v| <= tol (count_zeros); false: |v| > tol
    __host__ __device__ uint64_t operator()(uint64_t i) const {
        if (i >= nnz) return 0;
        const double2 v = d[i];
        const double n = hypot(v.x, v.y);
        return keep_small ? uint64_t(n <= tol) : uint64_t(n > tol);
    }
};

using FlagIter = rocprim::transform_iterator<rocprim::counting_iterator<uint64_t>, KeepFlag, uint64_t>;

__global__ __launch_bounds__(256) void k_filter_scatter(uint64_t nnz, const double2 *__restrict__ d,
                                                        const uint64_t *__restrict__ ix,
                                                        const uint64_t *__restrict__ pos,
                                                        uint64_t *__restrict__ ixo,
                                                        double2 *__restrict__ dout) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= nnz) return;
    const uint64_t p = pos[i];
    if (pos[i + 1] != p) {  // kept
        ixo[p] = ix[i];
        dout[p] = d[i];
    }
}

__global__ __launch_bounds__(256) void k_filter_indptr(uint64_t nrows, uint64_t nnz,
                                                       const uint64_t *__restrict__ ip,
                                                       const uint64_t *__restrict__ pos,
                                                       uint64_t *__restrict__ ipo) {
    const uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r > nrows) return;
    const uint64_t e = ip[r] < nnz ? ip[r] : nnz;
    ipo[r] = pos[e];
}

static unsigned blocks_for(uint64_t n) { return unsigned((n + 255) / 256); }

int launch_count_zeros(uint64_t nnz, const double *d, double tol, uint64_t *count, hipStream_t s) {
    *count = 0;
    if (nnz == 0) return QR_OK;
    FlagIter it(rocprim::counting_iterator<uint64_t>(0),
                KeepFlag{reinterpret_cast<const double2 *>(d), nnz, tol, true});
    size_t tmp_bytes = 0;
    QR_HIP_CHECK(hipcub::DeviceReduce::Sum(nullptr, tmp_bytes, it, (uint64_t *)nullptr, nnz, s));
    void *tmp = nullptr;
    QR_HIP_CHECK(hipMallocAsync(&tmp, 256 + tmp_bytes, s));
    uint64_t *d_count = static_cast<uint64_t *>(tmp);
    hipError_t e = hipcub::DeviceReduce::Sum(static_cast<unsigned char *>(tmp) + 256, tmp_bytes, it,
                                             d_count, nnz, s);
    if (e == hipSuccess) e = hipMemcpyAsync(count, d_count, sizeof(uint64_t), hipMemcpyDeviceToHost, s);
    (void)hipFreeAsync(tmp, s);
    if (e != hipSuccess) return hip_fail(e, "count_zeros");
    QR_HIP_CHECK(hipStreamSynchronize(s));
    return QR_OK;
}

int launch_eliminate_zeros(uint64_t nrows, uint64_t nnz, const uint64_t *ip, const uint64_t *ix,
                           const double *d, double tol, uint64_t *ipo, uint64_t *ixo, double *dout,
                           uint64_t *nnz_out, hipStream_t s) {
    // pos[0..nnz]: exclusive scan of the keep flags over nnz + 1 items (the
    // last flag is 0, so pos[nnz] is the kept total)
    FlagIter it(rocprim::counting_iterator<uint64_t>(0),
                KeepFlag{reinterpret_cast<const double2 *>(d), nnz, tol, false});
    size_t tmp_bytes = 0;
    QR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, it, (uint64_t *)nullptr,
                                                  nnz + 1, s));
    const size_t pbytes = ((nnz + 1) * sizeof(uint64_t) + 255) & ~size_t(255);
    void *tmp = nullptr;
    QR_HIP_CHECK(hipMallocAsync(&tmp, pbytes + tmp_bytes, s));
    uint64_t *pos = static_cast<uint64_t *>(tmp);
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(static_cast<unsigned char *>(tmp) + pbytes,
                                                    tmp_bytes, it, pos, nnz + 1, s);
    if (e == hipSuccess && nnz > 0) {
        hipLaunchKernelGGL(k_filter_scatter, dim3(blocks_for(nnz)), dim3(256), 0, s, nnz,
                           reinterpret_cast<const double2 *>(d), ix, pos, ixo,
                           reinterpret_cast<double2 *>(dout));
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_filter_indptr, dim3(blocks_for(nrows + 1)), dim3(256), 0, s, nrows, nnz,
                           ip, pos, ipo);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(nnz_out, pos + nnz, sizeof(uint64_t), hipMemcpyDeviceToHost, s);
    (void)hipFreeAsync(tmp, s);
    if (e != hipSuccess) return hip_fail(e, "eliminate_zeros");
    QR_HIP_CHECK(hipStreamSynchronize(s));
    return QR_OK;
}

}  // namespace qr
