// qr_apply.hip -- matrix-free Pauli-sum apply and <x|H|x> for gfx950.
//
// y_v[r] = sum_g D_g(r) * x_v[r ^ x_g],  D_g(r) = sum_{t in g} (-1)^popcount(r&z_t) c'_t
//
// This equals the rowwise CSR build followed by the CSR matvec
// (qrusty/src/accel.rs:171-210, 338-370) without materialising the 2^n*G
// entries: D_g(r) is the CSR value of group g in row r (same term-order fold),
// and the gather x[r ^ x_g] is the CSR column.  One thread owns one row; the
// group/term tables are read with wave-uniform (scalar) loads; D_g(r) is
// computed once per row and reused for every state vector of the launch.
//
// Precision modes
//   FP64 : complex128 state, FP64 fused multiply-adds.
//   MIXED: complex64 state.  D_g(r) is merged in FP64 (the cancellation between
//          terms sharing an x mask happens there) and rounded to FP32; the
//          product with x and the sum over groups run in FP32, the sum with
//          TwoSum compensation.  <x|y> is accumulated in FP64 from the
//          uncompensated-sum pair (s + c, exact in FP64).
// <x|y> is reduced deterministically: per-thread, then per block (fixed tree),
// then one final pass over the block partials.
#include "qr_internal.h"

#include <type_traits>

namespace qr {

constexpr int kApplyThreads = 256;

template <bool kNarrow>
__device__ __forceinline__ uint32_t parity_of(uint64_t a, uint64_t b) {
    if (kNarrow) return __popc(uint32_t(a) & uint32_t(b)) & 1u;
    return __popcll(a & b) & 1u;
}

// Knuth's TwoSum: s + e == a + b exactly (round-to-nearest, no reassociation).
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bp = s - a;
    const float ap = s - bp;
    e = (a - ap) + (b - bp);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off /= 2) v += __shfl_xor(v, off, 64);
    return v;
}

// Reduces NV complex values over the block; thread 0 writes them to out[2*v].
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&er)[NV], double (&ei)[NV],
                                                   double *out) {
    __shared__ double red[kApplyThreads / 64][NV][2];
    const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double a = wave_sum(er[v]), b = wave_sum(ei[v]);
        if (lane == 0) { red[wid][v][0] = a; red[wid][v][1] = b; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double a = 0.0, b = 0.0;
            for (int w = 0; w < kApplyThreads / 64; ++w) { a += red[w][v][0]; b += red[w][v][1]; }
            out[2 * v] = a;
            out[2 * v + 1] = b;
        }
    }
}

template <int NV, bool kMixed, bool kNarrow>
__global__ __launch_bounds__(kApplyThreads) void k_apply(
    uint64_t N, uint32_t G, uint64_t rows_per_block, uint32_t nblocks_xcd,
    const uint64_t *__restrict__ gx, const uint32_t *__restrict__ gterm,
    const uint64_t *__restrict__ tz, const double2 *__restrict__ tc,
    const uint8_t *__restrict__ gkind, const void *__restrict__ xin, void *__restrict__ yout,
    double *__restrict__ partials, uint32_t nvec_total, uint32_t v0) {
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    const cT *__restrict__ x = reinterpret_cast<const cT *>(xin) + uint64_t(v0) * N;
    cT *__restrict__ y = yout ? reinterpret_cast<cT *>(yout) + uint64_t(v0) * N : nullptr;

    // blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
    // range of rows so partner rows r ^ x_g of nearby rows share its L2
    uint32_t b = blockIdx.x;
    if (nblocks_xcd) b = (b % 8) * nblocks_xcd + b / 8;
    const uint64_t row0 = uint64_t(b) * rows_per_block;
    const uint64_t row1 = row0 + rows_per_block < N ? row0 + rows_per_block : N;

    double er[NV], ei[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { er[v] = 0.0; ei[v] = 0.0; }

    for (uint64_t r = row0 + threadIdx.x; r < row1; r += kApplyThreads) {
        double ar[NV], ai[NV];
        float sr[NV], si[NV], cr[NV], ci[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            ar[v] = ai[v] = 0.0;
            sr[v] = si[v] = cr[v] = ci[v] = 0.0f;
        }
        for (uint32_t g = 0; g < G; ++g) {
            const uint32_t t0 = gterm[g], t1 = gterm[g + 1];
            double2 c = tc[t0];
            if (parity_of<kNarrow>(r, tz[t0])) { c.x = -c.x; c.y = -c.y; }
            double dr = c.x, di = c.y;
            for (uint32_t t = t0 + 1; t < t1; ++t) {
                c = tc[t];
                if (parity_of<kNarrow>(r, tz[t])) { c.x = -c.x; c.y = -c.y; }
                dr = dr + c.x;
                di = di + c.y;
            }
            const uint64_t col = r ^ gx[g];
            const uint8_t kind = gkind[g];
            if (!kMixed) {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const double2 b2 = reinterpret_cast<const double2 *>(x)[uint64_t(v) * N + col];
                    if (kind == 1) {
                        ar[v] = fma(dr, b2.x, ar[v]);
                        ai[v] = fma(dr, b2.y, ai[v]);
                    } else {
                        ar[v] = fma(dr, b2.x, ar[v]);
                        ar[v] = fma(-di, b2.y, ar[v]);
                        ai[v] = fma(dr, b2.y, ai[v]);
                        ai[v] = fma(di, b2.x, ai[v]);
                    }
                }
            } else {
                const float fr = float(dr), fi = float(di);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const float2 b2 = reinterpret_cast<const float2 *>(x)[uint64_t(v) * N + col];
                    float pr, pi;
                    if (kind == 1) {
                        pr = fr * b2.x;
                        pi = fr * b2.y;
                    } else {
                        pr = fmaf(fr, b2.x, -fi * b2.y);
                        pi = fmaf(fr, b2.y, fi * b2.x);
                    }
                    float e;
                    two_sum(sr[v], pr, sr[v], e);
                    cr[v] += e;
                    two_sum(si[v], pi, si[v], e);
                    ci[v] += e;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double yr, yi;
            if (!kMixed) {
                yr = ar[v];
                yi = ai[v];
                if (y) reinterpret_cast<double2 *>(y)[uint64_t(v) * N + r] = make_double2(yr, yi);
                const double2 xr = reinterpret_cast<const double2 *>(x)[uint64_t(v) * N + r];
                if (partials) {
                    er[v] = fma(xr.x, yr, er[v]); er[v] = fma(xr.y, yi, er[v]);
                    ei[v] = fma(xr.x, yi, ei[v]); ei[v] = fma(-xr.y, yr, ei[v]);
                }
            } else {
                yr = double(sr[v]) + double(cr[v]);
                yi = double(si[v]) + double(ci[v]);
                if (y) reinterpret_cast<float2 *>(y)[uint64_t(v) * N + r] =
                           make_float2(sr[v] + cr[v], si[v] + ci[v]);
                const float2 xr = reinterpret_cast<const float2 *>(x)[uint64_t(v) * N + r];
                if (partials) {
                    const double xa = xr.x, xb = xr.y;
                    er[v] = fma(xa, yr, er[v]); er[v] = fma(xb, yi, er[v]);
                    ei[v] = fma(xa, yi, ei[v]); ei[v] = fma(-xb, yr, ei[v]);
                }
            }
        }
    }
    if (partials)
        block_reduce_store<NV>(er, ei, partials + (uint64_t(blockIdx.x) * nvec_total + v0) * 2);
}

// out[2v], out[2v+1] = sum over blocks of partials[(blk*nvec + v)*2 + {0,1}]
__global__ __launch_bounds__(256) void k_reduce_partials(uint32_t nblocks, uint32_t nvec,
                                                         const double *__restrict__ partials,
                                                         double *__restrict__ out) {
    const uint32_t v = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (uint32_t i = threadIdx.x; i < nblocks; i += 256) {
        a += partials[(uint64_t(i) * nvec + v) * 2];
        b += partials[(uint64_t(i) * nvec + v) * 2 + 1];
    }
    __shared__ double ra[256], rb[256];
    ra[threadIdx.x] = a;
    rb[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s /= 2) {
        if (int(threadIdx.x) < s) {
            ra[threadIdx.x] += ra[threadIdx.x + s];
            rb[threadIdx.x] += rb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * v] = ra[0]; out[2 * v + 1] = rb[0]; }
}

template <bool kMixed, bool kNarrow>
static void launch_chunk(int nv, dim3 grid, hipStream_t s, uint64_t N, uint32_t G,
                         uint64_t rpb, uint32_t nbx, const DeviceTables *t, const void *x, void *y,
                         double *partials, uint32_t nvec, uint32_t v0) {
#define QR_APPLY_CASE(NVV)                                                                  \
    case NVV:                                                                               \
        hipLaunchKernelGGL((k_apply<NVV, kMixed, kNarrow>), grid, dim3(kApplyThreads), 0, s, \
                           N, G, rpb, nbx, t->gx, t->gterm, t->tz, t->tc, t->gkind, x, y,   \
                           partials, nvec, v0);                                             \
        break;
    switch (nv) {
        QR_APPLY_CASE(1)
        QR_APPLY_CASE(2)
        QR_APPLY_CASE(4)
        default: QR_APPLY_CASE(8)
    }
#undef QR_APPLY_CASE
}

int launch_pauli_apply(const qr_op *op, DeviceTables *t, int32_t precision, uint32_t nvec,
                       const void *d_x, void *d_y, double *d_expect, hipStream_t s) {
    const uint64_t N = uint64_t(1) << op->nq;
    const uint32_t G = uint32_t(op->G);
    // about 2048 blocks (8 per CU); rows per block a multiple of the block size
    uint64_t rpb = (N + 2047) / 2048;
    rpb = (rpb + kApplyThreads - 1) / kApplyThreads * kApplyThreads;
    const uint64_t nblocks = (N + rpb - 1) / rpb;
    const uint32_t nbx = (nblocks % 8 == 0) ? uint32_t(nblocks / 8) : 0;
    double *partials = nullptr;
    if (d_expect) {
        const size_t need = size_t(nblocks) * nvec * 2;
        if (t->partials_cap < need) {
            if (t->partials) (void)hipFree(t->partials);
            t->partials = nullptr;
            t->partials_cap = 0;
            QR_HIP_CHECK(hipMalloc(&t->partials, need * sizeof(double)));
            t->partials_cap = need;
        }
        partials = t->partials;
    }
    const bool mixed = precision == QR_PREC_MIXED, narrow = op->nq <= 32;
    for (uint32_t v0 = 0; v0 < nvec;) {
        const uint32_t left = nvec - v0;
        const int nv = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
        const void *x = d_x;
        void *y = d_y;
        const dim3 grid(static_cast<unsigned>(nblocks));
        if (mixed) {
            if (narrow) launch_chunk<true, true>(nv, grid, s, N, G, rpb, nbx, t, x, y, partials, nvec, v0);
            else launch_chunk<true, false>(nv, grid, s, N, G, rpb, nbx, t, x, y, partials, nvec, v0);
        } else {
            if (narrow) launch_chunk<false, true>(nv, grid, s, N, G, rpb, nbx, t, x, y, partials, nvec, v0);
            else launch_chunk<false, false>(nv, grid, s, N, G, rpb, nbx, t, x, y, partials, nvec, v0);
        }
        QR_HIP_CHECK(hipGetLastError());
        v0 += uint32_t(nv);
    }
    if (d_expect) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(nvec), dim3(256), 0, s, uint32_t(nblocks), nvec,
                           partials, d_expect);
        QR_HIP_CHECK(hipGetLastError());
    }
    return QR_OK;
}

}  // namespace qr
