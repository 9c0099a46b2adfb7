// qr_apply.hip -- matrix-free Pauli-sum apply and <x|H|x> for gfx950.
//
// y_v[r] = sum_g D_g(r) * x_v[r ^ x_g],  D_g(r) = sum_{t in g} (-1)^popcount(r&z_t) c'_t
//
// This equals the rowwise CSR build followed by the CSR matvec
// (qrusty/src/accel.rs:171-210, 338-370) without materialising the 2^n*G
// entries: D_g(r) is the CSR value of group g in row r, x[r ^ x_g] the gather
// of its column.
//
// Work decomposition (one block = one tile of 2^L consecutive rows of one
// state vector):
//  * The tile's slice of x is staged in LDS (coalesced); groups whose x mask
//    is below 2^L gather from LDS, the others ("far" groups) gather the
//    partner tile t ^ (x_g >> L) from global memory (coalesced, permuted
//    inside a line).
//  * Blocks are ordered vector-major, and each XCD receives one contiguous
//    range of tiles of the vector: the far partners of nearby tiles are read
//    through the same L2, and one complex64 vector (128 MiB at 24 qubits)
//    stays resident in the 256 MB MALL while the chip works on it.
//  * Row r = h + l (h = tile base, l < 2^L):  (-1)^popcount(r & z) =
//    (-1)^popcount(h & z) * (-1)^popcount(l & zl).  The host merges the terms
//    of a group that share zl into "patterns"; per tile each pattern's
//    coefficient A_p = sum (-1)^popcount(h & z_t) c'_t is formed once (wave-
//    uniform), and per row only patterns with zl != 0 cost a sign.  The block
//    forms all its patterns' coefficients in parallel at the start of the
//    tile and keeps them in LDS next to the x tile.  Thread `tid` owns rows
//    l = tid + 1024*k: the sign splits again into a per-thread part (bits of
//    tid) and a per-k part (wave-uniform), so each pattern costs one FMA per
//    row.
//
// Precision modes
//   FP64 : complex128 state, FP64 FMAs (tile 2^12 rows, 4 rows per thread).
//   MIXED: complex64 state (tile 2^13 rows, 8 rows per thread).  D_g(r) is
//          merged in FP64 -- where terms sharing an x mask cancel -- and then
//          rounded to FP32.  Products and the sum over groups run in FP32:
//          products are accumulated into a block partial over up to
//          kCompBlock groups, and block partials are added to the running sum
//          with TwoSum (s + e == a + b exactly), the error terms collected in a
//          separate compensation.  <x|y> is reduced in FP64 from s + c.
// <x|y> is reduced deterministically: per thread, per block (fixed tree),
// then one pass over the block partials of each vector.
#include "qr_internal.h"

#include <type_traits>

namespace qr {

constexpr int kApplyThreads = 1024;
constexpr int kApplyThreadBits = 10;
constexpr int kCompBlock = 13;      // groups per FP32 partial between TwoSum steps
constexpr uint32_t kMaxPatLDS = 4096;  // per-tile pattern coefficients kept in LDS

template <bool kMixed>
struct ApplyShape {
    static constexpr int L = kMixed ? 13 : 12;           // tile = 2^L rows
    static constexpr int R = (1 << L) / kApplyThreads;  // rows per thread
};

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

// One group for the R rows of a thread: D_g(r) from the tile's pattern
// coefficients, then the gather of x[r ^ x_g] and the product into the
// accumulators (FP64: ar/ai; mixed: the FP32 block partial br/bi).
// With the coefficients in LDS (kPatLDS) the patterns are taken by class:
// thread-bit patterns once per thread, row-bit patterns from the per-tile
// table sU, straddling patterns per row.  Otherwise every pattern is formed
// on the fly and applied per row.
template <bool kMixed, bool kReal, bool kFull, bool kPatLDS, int R, typename cT>
__device__ __forceinline__ void apply_group(
    uint32_t g, uint32_t tid, uint64_t tile_rows, uint64_t base, const uint64_t *__restrict__ gx,
    const uint32_t *__restrict__ gpat, const uint32_t *__restrict__ gcls,
    const int32_t *__restrict__ gU, const uint32_t *__restrict__ pzl,
    const uint32_t *__restrict__ ptptr, const uint64_t *__restrict__ ptz,
    const double2 *__restrict__ ptc, const double2 *sA, const double2 *sU, const cT *xt,
    const cT *__restrict__ xv, double (&ar)[R], double (&ai)[R], float (&br)[R],
    float (&bi)[R]) {
    const uint64_t m = gx[g];
    double dr[R], di[R];
    const uint32_t p0 = gpat[g], p1 = gpat[g + 1];
    if (kPatLDS) {
        const uint32_t ps = gcls[2 * g], pk = gcls[2 * g + 1];
        // thread-bit patterns: one sign per thread
        double cr = 0.0, ci = 0.0;
        for (uint32_t p = p0; p < ps; ++p) {
            const double2 a = sA[p];
            const double sgn = (__popc(tid & pzl[p]) & 1u) ? -1.0 : 1.0;
            cr = fma(sgn, a.x, cr);
            if (!kReal) ci = fma(sgn, a.y, ci);
        }
        const int32_t u = gU[g];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            dr[k] = cr;
            di[k] = ci;
            if (u >= 0) {  // row-bit patterns: per-tile table
                const double2 uk = sU[u * R + k];
                dr[k] += uk.x;
                if (!kReal) di[k] += uk.y;
            }
        }
        // straddling patterns: per row
        for (uint32_t p = ps; p < pk; ++p) {
            const double2 a = sA[p];
            const uint32_t zl = pzl[p];
            const uint32_t st = __popc(tid & zl) & 1u;
            const double vr = st ? -a.x : a.x, vi = st ? -a.y : a.y;
            const uint32_t zk = zl >> kApplyThreadBits;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double sg = (__popc(uint32_t(k) & zk) & 1u) ? -1.0 : 1.0;  // uniform
                dr[k] = fma(sg, vr, dr[k]);
                if (!kReal) di[k] = fma(sg, vi, di[k]);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) dr[k] = di[k] = 0.0;
        for (uint32_t p = p0; p < p1; ++p) {
            double a_r = 0.0, a_i = 0.0;
            const uint32_t t1 = ptptr[p + 1];
            for (uint32_t t = ptptr[p]; t < t1; ++t) {
                const double2 c = ptc[t];
                if (__popcll(base & ptz[t]) & 1) { a_r -= c.x; a_i -= c.y; }
                else { a_r += c.x; a_i += c.y; }
            }
            // zl == 0 gives sign +1 everywhere and fma(1, a, d) == d + a
            const uint32_t zl = pzl[p];
            const uint32_t st = __popc(tid & zl) & 1u;  // per-thread sign
            const double vr = st ? -a_r : a_r, vi = st ? -a_i : a_i;
            const uint32_t zk = zl >> kApplyThreadBits;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double sg = (__popc(uint32_t(k) & zk) & 1u) ? -1.0 : 1.0;  // uniform
                dr[k] = fma(sg, vr, dr[k]);
                if (!kReal) di[k] = fma(sg, vi, di[k]);
            }
        }
    }
    // gather partner rows: l ^ (m & (2^L-1)) inside tile base ^ (m >> L)
    const uint64_t mlo = m & (tile_rows - 1), mhi = m & ~(tile_rows - 1);
    const uint32_t tx = tid ^ uint32_t(mlo & (kApplyThreads - 1));
    const uint32_t mk = uint32_t(mlo >> kApplyThreadBits);
    const cT *__restrict__ far = xv + (base ^ mhi);
    cT xg[R];
    if (mhi) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (kFull || tid + uint64_t(k) * kApplyThreads < tile_rows)
                xg[k] = far[tx + ((uint32_t(k) ^ mk) << kApplyThreadBits)];
        }
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (kFull || tid + uint64_t(k) * kApplyThreads < tile_rows)
                xg[k] = xt[tx + ((uint32_t(k) ^ mk) << kApplyThreadBits)];
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
        const cT xb = xg[k];
        if (!kMixed) {
            ar[k] = fma(dr[k], xb.x, ar[k]);
            ai[k] = fma(dr[k], xb.y, ai[k]);
            if (!kReal) {
                ar[k] = fma(-di[k], xb.y, ar[k]);
                ai[k] = fma(di[k], xb.x, ai[k]);
            }
        } else {
            const float fr = float(dr[k]);
            br[k] = fmaf(fr, xb.x, br[k]);
            bi[k] = fmaf(fr, xb.y, bi[k]);
            if (!kReal) {
                const float fi = float(di[k]);
                br[k] = fmaf(-fi, xb.y, br[k]);
                bi[k] = fmaf(fi, xb.x, bi[k]);
            }
        }
    }
}

template <bool kMixed, bool kNarrow, bool kFull, bool kPatLDS>
__global__ __launch_bounds__(kApplyThreads) void k_apply_tiled(
    uint64_t N, uint32_t G, uint32_t P, uint32_t ntiles, uint32_t tiles_per_xcd,
    const uint64_t *__restrict__ gx, const uint8_t *__restrict__ gkind,
    const uint32_t *__restrict__ gpat, const uint32_t *__restrict__ gcls,
    const int32_t *__restrict__ gU, const uint32_t *__restrict__ ug, uint32_t nU,
    const uint32_t *__restrict__ pzl, const uint32_t *__restrict__ ptptr,
    const uint64_t *__restrict__ ptz, const double2 *__restrict__ ptc,
    const void *__restrict__ xin, void *__restrict__ yout, double *__restrict__ partials) {
    using Shape = ApplyShape<kMixed>;
    constexpr int L = Shape::L, R = Shape::R;
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cT *xt = reinterpret_cast<cT *>(smem_raw);

    const uint32_t tid = threadIdx.x;
    const uint32_t v = blockIdx.x / ntiles;
    uint32_t tile = blockIdx.x % ntiles;
    if (tiles_per_xcd) tile = (tile % 8) * tiles_per_xcd + tile / 8;
    const uint64_t tile_rows = kFull ? (uint64_t(1) << L) : (N < (uint64_t(1) << L) ? N : (uint64_t(1) << L));
    const uint64_t base = uint64_t(tile) * tile_rows;
    const cT *__restrict__ xv = reinterpret_cast<const cT *>(xin) + uint64_t(v) * N;

    for (uint64_t i = tid; i < tile_rows; i += kApplyThreads) xt[i] = xv[base + i];
    // per-tile pattern coefficients A_p = sum_t (-1)^popcount(base & z_t) c'_t
    double2 *sA = reinterpret_cast<double2 *>(smem_raw + tile_rows * sizeof(cT));
    if (kPatLDS) {
        for (uint32_t p = tid; p < P; p += kApplyThreads) {
            double a_r = 0.0, a_i = 0.0;
            const uint32_t t1 = ptptr[p + 1];
            for (uint32_t t = ptptr[p]; t < t1; ++t) {
                const double2 c = ptc[t];
                if (parity_of<kNarrow>(base, ptz[t])) { a_r -= c.x; a_i -= c.y; }
                else { a_r += c.x; a_i += c.y; }
            }
            sA[p] = make_double2(a_r, a_i);
        }
    }
    __syncthreads();
    // per-tile row-bit table: sU[u*R + k] = sum over the row-bit patterns of
    // group ug[u] of (-1)^popcount(k & (zl >> thread bits)) * A_p
    double2 *sU = sA + P;
    if (kPatLDS && nU) {
        for (uint32_t i = tid; i < nU * R; i += kApplyThreads) {
            const uint32_t u = i / R, k = i % R, gg = ug[u];
            double s_r = 0.0, s_i = 0.0;
            for (uint32_t p = gcls[2 * gg + 1]; p < gpat[gg + 1]; ++p) {
                const double2 a = sA[p];
                if (__popc(k & (pzl[p] >> kApplyThreadBits)) & 1u) { s_r -= a.x; s_i -= a.y; }
                else { s_r += a.x; s_i += a.y; }
            }
            sU[i] = make_double2(s_r, s_i);
        }
        __syncthreads();
    }

    // accumulators: FP64 (ar, ai) or FP32 running sum s, compensation c, block partial b
    double ar[R], ai[R];
    float sr[R], si[R], cr[R], ci[R], br[R], bi[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        ar[k] = ai[k] = 0.0;
        sr[k] = si[k] = cr[k] = ci[k] = br[k] = bi[k] = 0.0f;
    }

    int since_flush = 0;
    for (uint32_t g = 0; g < G; ++g) {
        // kind 1: every c' of the group is real, so D_g(r) is real
        if (gkind[g] == 1)
            apply_group<kMixed, true, kFull, kPatLDS, R>(g, tid, tile_rows, base, gx, gpat, gcls,
                                                         gU, pzl, ptptr, ptz, ptc, sA, sU, xt, xv,
                                                         ar, ai, br, bi);
        else
            apply_group<kMixed, false, kFull, kPatLDS, R>(g, tid, tile_rows, base, gx, gpat, gcls,
                                                          gU, pzl, ptptr, ptz, ptc, sA, sU, xt, xv,
                                                          ar, ai, br, bi);
        if (kMixed && (++since_flush == kCompBlock || g + 1 == G)) {
            since_flush = 0;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                float e;
                two_sum(sr[k], br[k], sr[k], e);
                cr[k] += e;
                two_sum(si[k], bi[k], si[k], e);
                ci[k] += e;
                br[k] = bi[k] = 0.0f;
            }
        }
    }

    double er = 0.0, ei = 0.0;
    cT *__restrict__ yv = yout ? reinterpret_cast<cT *>(yout) + uint64_t(v) * N : nullptr;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint64_t l = tid + uint64_t(k) * kApplyThreads;
        if (!kFull && l >= tile_rows) break;
        double yr, yi;
        if (!kMixed) {
            yr = ar[k];
            yi = ai[k];
            if (yv) reinterpret_cast<double2 *>(yv)[base + l] = make_double2(yr, yi);
        } else {
            yr = double(sr[k]) + double(cr[k]);
            yi = double(si[k]) + double(ci[k]);
            if (yv) reinterpret_cast<float2 *>(yv)[base + l] = make_float2(sr[k] + cr[k], si[k] + ci[k]);
        }
        if (partials) {
            const cT xr = xt[l];
            const double xa = xr.x, xb = xr.y;
            er = fma(xa, yr, er); er = fma(xb, yi, er);
            ei = fma(xa, yi, ei); ei = fma(-xb, yr, ei);
        }
    }
    if (partials) {
        __shared__ double red[kApplyThreads / 64][2];
        const double a = wave_sum(er), b = wave_sum(ei);
        if ((tid & 63) == 0) { red[tid / 64][0] = a; red[tid / 64][1] = b; }
        __syncthreads();
        if (tid == 0) {
            double s0 = 0.0, s1 = 0.0;
            for (int w = 0; w < kApplyThreads / 64; ++w) { s0 += red[w][0]; s1 += red[w][1]; }
            partials[2 * uint64_t(blockIdx.x)] = s0;
            partials[2 * uint64_t(blockIdx.x) + 1] = s1;
        }
    }
}

// out[2v], out[2v+1] = sum of the ntiles block partials of vector v
__global__ __launch_bounds__(256) void k_reduce_partials(uint32_t ntiles,
                                                         const double *__restrict__ partials,
                                                         double *__restrict__ out) {
    const uint32_t v = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (uint32_t i = threadIdx.x; i < ntiles; i += 256) {
        a += partials[2 * (uint64_t(v) * ntiles + i)];
        b += partials[2 * (uint64_t(v) * ntiles + i) + 1];
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
static int launch_tiled(const qr_op *op, DeviceTables *t, uint32_t nvec, const void *d_x,
                        void *d_y, double *partials, uint32_t ntiles, hipStream_t s) {
    using Shape = ApplyShape<kMixed>;
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    const uint64_t N = uint64_t(1) << op->nq;
    const uint64_t tile_rows = N < (uint64_t(1) << Shape::L) ? N : (uint64_t(1) << Shape::L);
    const DeviceTables::Patterns &pt = t->pat[kMixed ? 1 : 0];
    const bool plds = pt.P + size_t(pt.nU) * Shape::R <= kMaxPatLDS;
    const size_t shmem = size_t(tile_rows) * sizeof(cT) +
                         (plds ? (size_t(pt.P) + size_t(pt.nU) * Shape::R) * sizeof(double2) : 0);
    const uint32_t tpx = (ntiles % 8 == 0) ? ntiles / 8 : 0;
    const bool full = tile_rows == (uint64_t(1) << Shape::L);
    auto kern = full ? (plds ? k_apply_tiled<kMixed, kNarrow, true, true>
                             : k_apply_tiled<kMixed, kNarrow, true, false>)
                     : (plds ? k_apply_tiled<kMixed, kNarrow, false, true>
                             : k_apply_tiled<kMixed, kNarrow, false, false>);
    QR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    const uint64_t nblocks = uint64_t(ntiles) * nvec;
    if (nblocks > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "apply grid too large");
    hipLaunchKernelGGL(kern, dim3(unsigned(nblocks)), dim3(kApplyThreads), shmem, s, N,
                       uint32_t(op->G), pt.P, ntiles, tpx, t->gx, t->gkind, pt.gpat, pt.gcls, pt.gU,
                       pt.ug, pt.nU, pt.pzl, pt.ptptr, pt.ptz, pt.ptc, d_x, d_y, partials);
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

int launch_pauli_apply(const qr_op *op, DeviceTables *t, int32_t precision, uint32_t nvec,
                       const void *d_x, void *d_y, double *d_expect, hipStream_t s) {
    const bool mixed = precision == QR_PREC_MIXED, narrow = op->nq <= 32;
    const int Lmax = mixed ? ApplyShape<true>::L : ApplyShape<false>::L;
    const int L = int(op->nq) < Lmax ? int(op->nq) : Lmax;
    int rc = device_patterns(op, t, mixed ? 1 : 0, L, kApplyThreadBits);
    if (rc) return rc;
    const uint64_t N = uint64_t(1) << op->nq;
    const uint32_t ntiles = uint32_t(N >> L);
    double *partials = nullptr;
    if (d_expect) {
        const size_t need = size_t(ntiles) * nvec * 2;
        if (t->partials_cap < need) {
            if (t->partials) (void)hipFree(t->partials);
            t->partials = nullptr;
            t->partials_cap = 0;
            QR_HIP_CHECK(hipMalloc(&t->partials, need * sizeof(double)));
            t->partials_cap = need;
        }
        partials = t->partials;
    }
    if (mixed) rc = narrow ? launch_tiled<true, true>(op, t, nvec, d_x, d_y, partials, ntiles, s)
                           : launch_tiled<true, false>(op, t, nvec, d_x, d_y, partials, ntiles, s);
    else rc = narrow ? launch_tiled<false, true>(op, t, nvec, d_x, d_y, partials, ntiles, s)
                     : launch_tiled<false, false>(op, t, nvec, d_x, d_y, partials, ntiles, s);
    if (rc) return rc;
    if (d_expect) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(nvec), dim3(256), 0, s, ntiles, partials,
                           d_expect);
        QR_HIP_CHECK(hipGetLastError());
    }
    return QR_OK;
}

}  // namespace qr
