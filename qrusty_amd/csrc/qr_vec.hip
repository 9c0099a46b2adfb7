// qr_vec.hip -- the eigensolver-step operations around the matvec, on the
// device: the CSR diagonal, a x + b y and the diagonal preconditioner.
//
// They replace pyqrusty's SpMat.diagonal (pyqrusty/src/lib.rs:118-125, sprs
// CsMat::diag().to_dense()), qrusty::accel::axpby / axpy / ax
// (qrusty/src/accel.rs:374-393) and precond / precond2 / reg
// (pyqrusty/src/lib.rs:436-468).  Complex arithmetic is num-complex's:
// products (a.re x.re - a.im x.im, a.re x.im + a.im x.re) and sums with no
// fused multiply-add (the library is built with -ffp-contract=off), division
// ((a c + b d) / (c c + d d), (b c - a d) / (c c + d d)), norm = hypot.
// Every kernel is a grid-stride loop over its elements.
#include "qr_internal.h"

namespace qr {

__device__ __forceinline__ double2 cmul(double2 a, double2 x) {
    return make_double2(a.x * x.x - a.y * x.y, a.x * x.y + a.y * x.x);
}

__device__ __forceinline__ double2 cdiv(double2 p, double2 q) {
    const double n = q.x * q.x + q.y * q.y;
    const double re = p.x * q.x + p.y * q.y;
    const double im = p.y * q.x - p.x * q.y;
    return make_double2(re / n, im / n);
}

// Diagonal entry of row r: the entry with column r, found by binary search in
// the row's ascending columns (as sprs CsVecView::get); 0 when absent.
__device__ __forceinline__ double2 diag_entry(uint64_t r, uint64_t nnz,
                                              const uint64_t *__restrict__ indptr,
                                              const uint64_t *__restrict__ indices,
                                              const double2 *__restrict__ data) {
    uint64_t lo = indptr[r], hi = indptr[r + 1];
    if (hi > nnz) hi = nnz;
    if (lo > hi) lo = hi;
    while (lo < hi) {  // first position with column >= r
        const uint64_t mid = lo + (hi - lo) / 2;
        if (indices[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    if (lo < (indptr[r + 1] < nnz ? indptr[r + 1] : nnz) && indices[lo] == r) return data[lo];
    return make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_csr_diagonal(uint64_t n, uint64_t nnz,
                                                      const uint64_t *__restrict__ indptr,
                                                      const uint64_t *__restrict__ indices,
                                                      const double2 *__restrict__ data,
                                                      double2 *__restrict__ out) {
    for (uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * 256)
        out[r] = diag_entry(r, nnz, indptr, indices, data);
}

// z = a x + b y  (mode 0), a x + y (mode 1), a x (mode 2)
__global__ __launch_bounds__(256) void k_axpby(uint64_t n, int mode, double2 a,
                                               const double2 *__restrict__ x, double2 b,
                                               const double2 *__restrict__ y,
                                               double2 *__restrict__ z) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * 256) {
        const double2 ax = cmul(a, x[i]);
        if (mode == 2) {
            z[i] = ax;
        } else {
            const double2 by = mode == 0 ? cmul(b, y[i]) : y[i];
            z[i] = make_double2(ax.x + by.x, ax.y + by.y);
        }
    }
}

// out = dx / reg(d - e, tol), reg(v) = tol + 0i when |v| < tol, else v; the
// diagonal d either given (precond2) or read from the CSR (precond)
__global__ __launch_bounds__(256) void k_precond(uint64_t n, const double2 *__restrict__ diag,
                                                 uint64_t nnz, const uint64_t *__restrict__ indptr,
                                                 const uint64_t *__restrict__ indices,
                                                 const double2 *__restrict__ data,
                                                 const double2 *__restrict__ dx, double2 e,
                                                 double tol, double2 *__restrict__ out) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * 256) {
        const double2 d = diag ? diag[i] : diag_entry(i, nnz, indptr, indices, data);
        double2 v = make_double2(d.x - e.x, d.y - e.y);
        if (hypot(v.x, v.y) < tol) v = make_double2(tol, 0.0);
        out[i] = cdiv(dx[i], v);
    }
}

static unsigned grid_for(uint64_t n) {
    const uint64_t b = (n + 255) / 256;
    return unsigned(b < 65536 ? (b ? b : 1) : 65536);
}

int launch_csr_diagonal(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *indptr,
                        const uint64_t *indices, const double *data, double *out, hipStream_t s) {
    const uint64_t n = nrows < ncols ? nrows : ncols;
    if (n == 0) return QR_OK;
    hipLaunchKernelGGL(k_csr_diagonal, dim3(grid_for(n)), dim3(256), 0, s, n, nnz, indptr,
                       indices, reinterpret_cast<const double2 *>(data),
                       reinterpret_cast<double2 *>(out));
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

int launch_axpby(uint64_t n, int mode, const double *a, const double *x, const double *b,
                 const double *y, double *z, hipStream_t s) {
    if (n == 0) return QR_OK;
    const double2 a2 = make_double2(a[0], a[1]);
    const double2 b2 = b ? make_double2(b[0], b[1]) : make_double2(0.0, 0.0);
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, n, mode, a2,
                       reinterpret_cast<const double2 *>(x), b2,
                       reinterpret_cast<const double2 *>(y), reinterpret_cast<double2 *>(z));
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

int launch_precond(uint64_t n, const double *diag, uint64_t nnz, const uint64_t *indptr,
                   const uint64_t *indices, const double *data, const double *dx, const double *e,
                   double tol, double *out, hipStream_t s) {
    if (n == 0) return QR_OK;
    hipLaunchKernelGGL(k_precond, dim3(grid_for(n)), dim3(256), 0, s, n,
                       reinterpret_cast<const double2 *>(diag), nnz, indptr, indices,
                       reinterpret_cast<const double2 *>(data),
                       reinterpret_cast<const double2 *>(dx), make_double2(e[0], e[1]), tol,
                       reinterpret_cast<double2 *>(out));
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

}  // namespace qr
