// qr_apply.hip -- matrix-free Pauli-sum apply and <x|H|x> for gfx950.
//
// y_v[r] = sum_g D_g(r) * x_v[r ^ x_g],  D_g(r) = sum_{t in g} (-1)^popcount(r&z_t) c'_t
//
// This equals the rowwise CSR build followed by the CSR matvec
// (qrusty/src/accel.rs:171-210, 338-370) without materialising the 2^n*G
// entries: D_g(r) is the CSR value of group g in row r, x[r ^ x_g] the gather
// of its column.
//
// Work decomposition (one block of 256 threads = one tile of 2^L consecutive
// rows of one state vector; four blocks share a CU, so that some compute
// while others wait at a barrier or on their loads):
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
//    l = tid + 256*k: the sign splits again into a per-thread part (bits of
//    tid) and a per-k part (wave-uniform), so each pattern costs one FMA per
//    row.  The gather form on full complex64 tiles owns row pairs instead,
//    l = 2*tid + (k & 1) + 512*(k >> 1), so that one 16-byte load gathers
//    both partners of a pair and one 16-byte store writes both results; the
//    patterns' low z bits are permuted to match (row bit 0 counts with k).
//  * Partner-staged form (k_apply_staged): when many far groups share a
//    partner tile (their masks agree above the tile), the groups are ordered
//    by partner and each partner tile is loaded into LDS once, double-
//    buffered, instead of being gathered once per group.
//
// Precision modes
//   FP64 : complex128 state, FP64 FMAs (tile 2^10 rows, 4 rows per thread).
//   MIXED: complex64 state (tile 2^11 rows, 8 rows per thread).  D_g(r) is
//          merged in FP64 -- where terms sharing an x mask cancel -- and then
//          rounded to FP32.  Products and the sum over groups run in FP32:
//          products are accumulated into a block partial over up to
//          kCompBlock groups, and block partials are added to the running sum
//          with TwoSum (s + e == a + b exactly); the error e starts the next
//          block partial, so y = s + b at the end.  <x|y> is reduced in FP64
//          from s + b.
// <x|y> is reduced deterministically: per thread, per block (fixed tree),
// then one pass over the block partials of each vector.
#include "qr_internal.h"

#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>

namespace qr {

// 256 threads and at most 40 KB of LDS per block: four blocks per CU (the
// 160 KB of a gfx950 CU).  Measured (24-qubit Heisenberg, 8 states; 28-qubit
// k-local, one state), by threads per block and tile:
//   1024, 2^13 / 2^12 rows (mixed / FP64): 1.80 / 3.75 ms; 252 / 422 ms
//    512, 2^12 / 2^11 rows:                1.75 / 3.66 ms; 237 / 398 ms
//    256, 2^11 / 2^10 rows:                1.64 / 3.47 ms; 228 / 400 ms
// although each halving of the tile adds a far group.  Fewer rows per thread
// on the same tiles (512 threads, 2^11 / 2^10 rows: 4 / 2 rows per thread,
// 72-88 VGPRs, 5-7 waves per SIMD) measured 1.84 / 3.59 ms and 278 / 436 ms
// against 1.63 / 3.50 and 238 / 401 with 256 threads (block partials summed
// in FP64 in both): the per-group work of a thread is shared by fewer rows.
// Forcing 5 or 6 waves per SIMD on 256 threads spills 120-330 registers.
constexpr int kApplyThreads = 256;
constexpr int kApplyThreadBits = 8;
// Groups per FP32 block partial between TwoSum steps.  Measured on the same
// 24-qubit Heisenberg case (8 states, psi rounded to complex64; error of
// <psi|H|psi> relative to the FP64 value on the rounded state, max over
// states), by block size: 1: 2.6e-7, 2.16 ms; 4: 2.8e-7, 1.69 ms;
// 16: 5.9e-7, 1.62 ms; 32: 6.6e-7, 1.62 ms.  4 keeps the whole-call error of
// <psi|H|psi> (rounding of psi included) near 5e-7.  Those figures are
// with the TwoSum errors collected in a separate compensation c; feeding
// each error into the next block partial instead saves the compensation's
// adds, zeroing and 16 registers (113 -> 95 VGPRs, 5 waves per SIMD): 1.651
// -> 1.583 ms, E 5.05e-7 -> 5.22e-7 end to end, y unchanged.  Adding the
// block partials into FP64 running sums gave 1.63 ms; the accumulation stays
// FP32 with TwoSum compensation, FP64 being kept to the merging of
// coefficients and the <x|y> reduction.
constexpr int kCompBlock = 4;
constexpr size_t kApplyMaxLDS = 40 * 1024;  // LDS of one block

// Tile of 2^L rows (x tile in LDS: 16 KB in both precisions at the default
// L; 8 KB for the small complex64 tiles), R rows per thread, and the LDS left
// for one chunk of pattern coefficients (20 B each) and the row-bit table
// (R x 16 B per row, plus a zero row).
//
// Small complex64 tiles (2^10 rows, 4 per thread) serve launches that would
// leave CUs short of blocks with 2^11-row tiles: the block keeps the latency
// of its LDS and gather round trips hidden only with others beside it.
// Measured, one state (20 qubits, 14251 terms in 2536 groups; 512 blocks of
// 2^11 rows, two per CU): 2.29 -> 1.70 ms.  Where the default tiles fill the
// chip they are 20-27% slower (22- and 24-qubit operators of 811-3792
// groups, the 20-qubit one with 4 states, 24-qubit Heisenberg with 8).
constexpr int kMixedSmallL = 10;
template <bool kMixed, int kL = kMixed ? 11 : 10>
struct ApplyShape {
    static constexpr int L = kL;
    static constexpr int R = (1 << L) / kApplyThreads;
    static constexpr size_t kTile = (size_t(1) << L) * (kMixed ? 8 : 16);
    // LDS after the x tile: 5/8 for patterns, 3/8 for the row-bit table
    static constexpr size_t kRest = kApplyMaxLDS - kTile - 1024;
    static constexpr uint32_t kPatChunk = uint32_t(kRest * 5 / 8 / 20) < 2048u ? uint32_t(kRest * 5 / 8 / 20) : 2048u;
    static constexpr uint32_t kTableRows = uint32_t(kRest * 3 / 8 / (R * 16)) - 1;
    // partner-staged form: two tile buffers, the rest split the same way
    static constexpr size_t kRestS = kApplyMaxLDS - 2 * kTile - 1024;
    static constexpr uint32_t kPatChunkS = uint32_t(kRestS * 5 / 8 / 20);
    static constexpr uint32_t kTableRowsS = uint32_t(kRestS * 3 / 8 / (R * 16)) - 1;
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

// x(-1) when bit 31 of `flip` is set: the sign of a double flipped by an XOR
// of its high word.
__device__ __forceinline__ double flip_sign(double a, uint32_t flip) {
    return __hiloint2double(__double2hiint(a) ^ int(flip), __double2loint(a));
}

// Word 3 of a raw (untyped, stride 0) buffer resource on gfx9: 32-bit data
// format; offsets are bounds-checked against num_records bytes.
constexpr int kBufferWord3 = 0x00020000;

// Gathers of x[r ^ m] for the R rows of a thread: from the LDS x tile, or --
// for masks above the tile -- from the partner tile in global memory, with
// one uniform base pointer per row block and the thread's offset shared.
// With kPairs (complex64, full tiles) the thread's rows k = 2j, 2j+1 are the
// adjacent rows (2 tid, 2 tid + 1) + 2^9 j: their partners r ^ m form one
// aligned pair, read with one 16-byte load.  When bit 0 of m is set, row k's
// partner is xg[k ^ 1]: such groups form classes of their own, whose
// accumulation reads the pair crosswise (no data is moved).
template <bool kFar, bool kFull, bool kPairs, int R, typename cT>
__device__ __forceinline__ void gather(uint64_t m, uint32_t tid, uint64_t tile_rows, uint64_t base,
                                       const cT *xt, const cT *__restrict__ xv, cT (&xg)[R]) {
    const uint64_t mlo = m & (tile_rows - 1), mhi = m & ~(tile_rows - 1);
    if constexpr (kPairs) {
        static_assert(sizeof(cT) == 8 && kFull && R % 2 == 0, "pairs: complex64, full tiles");
        constexpr int kJ = kApplyThreadBits + 1;  // row bit of the pair index j
        const uint32_t tq = tid ^ uint32_t((mlo >> 1) & (kApplyThreads - 1));
        const uint32_t jm = uint32_t(mlo >> kJ);
        float4 pr[R / 2];
        if (kFar) {
            const cT *far = xv + (base ^ mhi);
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<cT *>(far), 0, int(tile_rows * sizeof(cT)), kBufferWord3);
            const uint32_t voff = tq * uint32_t(2 * sizeof(cT));
#pragma unroll
            for (int j = 0; j < R / 2; ++j) {
                const uint32_t soff = ((uint32_t(j) ^ jm) << kJ) * uint32_t(sizeof(cT));
                pr[j] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
            }
        } else {
            const float4 *xrow = reinterpret_cast<const float4 *>(xt) + tq;
#pragma unroll
            for (int j = 0; j < R / 2; ++j) pr[j] = xrow[(uint32_t(j) ^ jm) << kApplyThreadBits];
        }
#pragma unroll
        for (int j = 0; j < R / 2; ++j) {
            xg[2 * j] = make_float2(pr[j].x, pr[j].y);
            xg[2 * j + 1] = make_float2(pr[j].z, pr[j].w);
        }
        return;
    }
    const uint32_t tx = tid ^ uint32_t(mlo & (kApplyThreads - 1));
    const uint32_t mk = uint32_t(mlo >> kApplyThreadBits);
    if (kFar) {
        // the partner tile through a buffer resource: the row block's offset
        // is a scalar, the thread's offset one VGPR shared by the R loads
        const cT *far = xv + (base ^ mhi);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<cT *>(far), 0, int(tile_rows * sizeof(cT)), kBufferWord3);
        const uint32_t voff = tx * uint32_t(sizeof(cT));
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
            const uint32_t soff = ((uint32_t(k) ^ mk) << kApplyThreadBits) * uint32_t(sizeof(cT));
            if constexpr (sizeof(cT) == 8)
                xg[k] = __builtin_bit_cast(cT, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
            else
                xg[k] = __builtin_bit_cast(cT, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
        }
    } else {
        // the thread's column in the LDS tile once, then a uniform row offset
        const cT *xrow = xt + tx;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
            xg[k] = xrow[(uint32_t(k) ^ mk) << kApplyThreadBits];
        }
    }
}

// y += D_g(r) x[r ^ m] for the R rows of a thread.  fetch() gathers x into
// xg (row k's partner in xg[k], or in xg[k ^ 1] with kSwap): first thing for
// groups with one D per thread, so that the loads fly during the pattern
// loop; after the per-row D for row-varying groups, so that the gathered
// registers and the coefficient temporaries are not live together.  D_g is
// formed from the chunk's pattern coefficients: thread-bit patterns once per
// thread, then -- for row-varying groups (kRowvar) -- per row the row-bit
// table entry sU[urow*R + k] and the straddling patterns.  Products go to the
// FP64 accumulators ar/ai, or (mixed) to the FP32 block partials br/bi; those
// with Im D only for complex groups (kCplx).
template <bool kMixed, bool kFull, bool kRowvar, bool kCplx, bool kSwap, int R, typename cT, typename Fetch>
__device__ __forceinline__ void accumulate(const ApplyGroup &gd, uint32_t tid, uint64_t tile_rows,
                                           uint32_t zero_row, const double2 *sA, const uint32_t *sZ,
                                           const double2 *sU, cT (&xg)[R], Fetch fetch, double (&ar)[R],
                                           double (&ai)[R], float (&br)[R], float (&bi)[R]) {
    constexpr bool cplx = kCplx;
    if (!kRowvar) fetch();
    double cr = 0.0, ci = 0.0;
    uint32_t p = gd.p0;
    for (; p + 1 < gd.ps; p += 2) {  // two patterns per step: loads in flight together
        const double2 a0 = sA[p], a1 = sA[p + 1];
        const uint32_t f0 = uint32_t(__popc(tid & sZ[p])) << 31;
        const uint32_t f1 = uint32_t(__popc(tid & sZ[p + 1])) << 31;
        cr += flip_sign(a0.x, f0);
        cr += flip_sign(a1.x, f1);
        if (cplx) {
            ci += flip_sign(a0.y, f0);
            ci += flip_sign(a1.y, f1);
        }
    }
    if (p < gd.ps) {
        const double2 a = sA[p];
        const uint32_t flip = uint32_t(__popc(tid & sZ[p])) << 31;
        cr += flip_sign(a.x, flip);
        if (cplx) ci += flip_sign(a.y, flip);
    }
    if (!kRowvar) {
        // D is the same for the R rows of this thread
        if (kMixed) {
            const float fr = float(cr), fi = float(ci);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
                br[k] = fmaf(fr, xg[k ^ int(kSwap)].x, br[k]);
                bi[k] = fmaf(fr, xg[k ^ int(kSwap)].y, bi[k]);
            }
            if (cplx) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
                    br[k] = fmaf(-fi, xg[k ^ int(kSwap)].y, br[k]);
                    bi[k] = fmaf(fi, xg[k ^ int(kSwap)].x, bi[k]);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
                ar[k] = fma(cr, xg[k ^ int(kSwap)].x, ar[k]);
                ai[k] = fma(cr, xg[k ^ int(kSwap)].y, ai[k]);
            }
            if (cplx) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
                    ar[k] = fma(-ci, xg[k ^ int(kSwap)].y, ar[k]);
                    ai[k] = fma(ci, xg[k ^ int(kSwap)].x, ai[k]);
                }
            }
        }
        return;
    }
    // D per row for the row-varying groups (rounded to FP32 in mixed mode),
    // formed by one of the paths below; a single block of products follows,
    // so that every path leaves the accumulators in the same registers
    using dT = typename std::conditional<kMixed, float, double>::type;
    dT dre[R], dim[R];
    if (!kFull) {
#pragma unroll
        for (int k = 0; k < R; ++k) dre[k] = dim[k] = dT(0);
    }
    const double2 *su = sU + (gd.u >= 0 ? uint32_t(gd.u) : zero_row) * R;
    const uint32_t nst = gd.pk - gd.ps;
    if (nst <= 2) {
        // at most two straddling patterns: each one's per-thread coefficient
        // (the sign of the thread bits applied) is formed once, and per row
        // only the wave-uniform sign of the row-index bits flips it
        auto rows = [&](auto nst_t) {
            constexpr uint32_t kN = decltype(nst_t)::value;
            double2 aq[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
            uint32_t zh[2] = {0u, 0u};
#pragma unroll
            for (uint32_t j = 0; j < kN; ++j) {
                const double2 a = sA[gd.ps + j];
                const uint32_t zl = uint32_t(__builtin_amdgcn_readfirstlane(int(sZ[gd.ps + j])));
                const uint32_t f = uint32_t(__popc(tid & zl)) << 31;
                aq[j] = make_double2(flip_sign(a.x, f), flip_sign(a.y, f));
                zh[j] = zl >> kApplyThreadBits;
            }
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
                const double2 uk = su[k];
                double dr = cr + uk.x, di = cplx ? ci + uk.y : 0.0;
#pragma unroll
                for (uint32_t j = 0; j < kN; ++j) {
                    // the row sign is wave-uniform: fma(+-1, a, d) rounds d +- a
                    // exactly as the add does, with the sign in scalar registers
                    const double sg = (__popc(uint32_t(k) & zh[j]) & 1u) ? -1.0 : 1.0;
                    dr = fma(sg, aq[j].x, dr);
                    if (cplx) di = fma(sg, aq[j].y, di);
                }
                dre[k] = dT(dr);
                if (cplx) dim[k] = dT(di);
            }
        };
        if (nst == 0) rows(std::integral_constant<uint32_t, 0>{});
        else if (nst == 1) rows(std::integral_constant<uint32_t, 1>{});
        else rows(std::integral_constant<uint32_t, 2>{});
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
            const double2 uk = su[k];
            double dr = cr + uk.x, di = cplx ? ci + uk.y : 0.0;
            for (uint32_t q = gd.ps; q < gd.pk; ++q) {
                const double2 a = sA[q];
                const uint32_t zl = uint32_t(__builtin_amdgcn_readfirstlane(int(sZ[q])));
                const uint32_t flip =
                    uint32_t(__popc(tid & zl) + __popc(uint32_t(k) & (zl >> kApplyThreadBits))) << 31;
                dr += flip_sign(a.x, flip);
                if (cplx) di += flip_sign(a.y, flip);
            }
            dre[k] = dT(dr);
            if (cplx) dim[k] = dT(di);
        }
    }
    fetch();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (!kFull && tid + uint64_t(k) * kApplyThreads >= tile_rows) break;
        const cT xb = xg[k ^ int(kSwap)];
        if (!kMixed) {
            ar[k] = fma(double(dre[k]), xb.x, ar[k]);
            ai[k] = fma(double(dre[k]), xb.y, ai[k]);
            if (cplx) {
                ar[k] = fma(-double(dim[k]), xb.y, ar[k]);
                ai[k] = fma(double(dim[k]), xb.x, ai[k]);
            }
        } else {
            br[k] = fmaf(float(dre[k]), xb.x, br[k]);
            bi[k] = fmaf(float(dre[k]), xb.y, bi[k]);
            if (cplx) {
                br[k] = fmaf(-float(dim[k]), xb.y, br[k]);
                bi[k] = fmaf(float(dim[k]), xb.x, bi[k]);
            }
        }
    }
}

// y is written once and not read again by the kernel: nontemporal stores, so
// they do not displace x (which the partner tiles still gather) from the caches.
template <typename cT>
__device__ __forceinline__ void store_y(cT *p, cT v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// A_p = sum_t (-1)^popcount(base & z_t) c'_t over the pattern's terms
// [t0, t1), in term order.  The loads of four terms are issued together; the
// sum itself stays sequential.
template <bool kNarrow>
__device__ __forceinline__ double2 pattern_coeff(uint64_t base, uint32_t t0, uint32_t t1,
                                                 const uint64_t *__restrict__ ptz,
                                                 const double2 *__restrict__ ptc) {
    double a_r = 0.0, a_i = 0.0;
    uint32_t t = t0;
    for (; t + 4 <= t1; t += 4) {
        double2 cc[4];
        uint32_t par[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cc[j] = ptc[t + j];
            par[j] = parity_of<kNarrow>(base, ptz[t + j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (par[j]) { a_r -= cc[j].x; a_i -= cc[j].y; }
            else { a_r += cc[j].x; a_i += cc[j].y; }
        }
    }
    for (; t < t1; ++t) {
        const double2 cc = ptc[t];
        if (parity_of<kNarrow>(base, ptz[t])) { a_r -= cc.x; a_i -= cc.y; }
        else { a_r += cc.x; a_i += cc.y; }
    }
    return make_double2(a_r, a_i);
}

// LDS layout of k_apply_tiled: x tile | sA[max_p] | sZ[max_p] | sU[max_u * R]
template <typename cT>
__host__ __device__ inline size_t apply_lds_offsets(uint64_t tile_rows, uint32_t max_p, uint32_t R,
                                                    uint32_t max_u, size_t *oA, size_t *oZ,
                                                    size_t *oU) {
    const size_t a = (size_t(tile_rows) * sizeof(cT) + 15) & ~size_t(15);
    const size_t z = a + size_t(max_p) * sizeof(double2);
    const size_t u = (z + size_t(max_p) * sizeof(uint32_t) + 15) & ~size_t(15);
    if (oA) *oA = a;
    if (oZ) *oZ = z;
    if (oU) *oU = u;
    return u + size_t(max_u) * R * sizeof(double2);
}

// Per-chunk LDS tables of a tile: the pattern coefficients
// sA[p] = A_p = sum_t (-1)^popcount(base & z_t) c'_t with their low z bits
// sZ[p], then the row-bit table sU[u*R + k] = sum over the row-bit patterns of
// the group of (-1)^popcount(k & (zl >> thread bits)) * A_p; row zero_row
// stays zero for groups without such patterns.  Ends with a barrier.
template <bool kNarrow, int R>
__device__ __forceinline__ void chunk_tables(const ApplyChunk &ch, uint32_t tid, uint64_t base,
                                             uint32_t zero_row, const ApplyGroup *__restrict__ grp,
                                             const uint32_t *__restrict__ ug,
                                             const uint32_t *__restrict__ pzl,
                                             const uint32_t *__restrict__ ptptr,
                                             const uint64_t *__restrict__ ptz,
                                             const double2 *__restrict__ ptc, double2 *sA,
                                             uint32_t *sZ, double2 *sU) {
    for (uint32_t p = tid; p < ch.p1 - ch.p0; p += kApplyThreads) {
        const uint32_t pa = ch.p0 + p;
        sA[p] = pattern_coeff<kNarrow>(base, ptptr[pa], ptptr[pa + 1], ptz, ptc);
        sZ[p] = pzl[pa];
    }
    __syncthreads();
    for (uint32_t i = tid; i < (ch.u1 - ch.u0 + 1) * R; i += kApplyThreads) {
        const uint32_t u = i / R, k = i % R;
        double s_r = 0.0, s_i = 0.0;
        if (u < ch.u1 - ch.u0) {
            const ApplyGroup &gg = grp[ug[ch.u0 + u]];
            for (uint32_t p = gg.pk; p < gg.p1; ++p) {
                const double2 a = sA[p];
                if (__popc(k & (sZ[p] >> kApplyThreadBits)) & 1u) { s_r -= a.x; s_i -= a.y; }
                else { s_r += a.x; s_i += a.y; }
            }
        }
        sU[(u < ch.u1 - ch.u0 ? u : zero_row) * R + k] = make_double2(s_r, s_i);
    }
    __syncthreads();
}

template <bool kMixed, bool kNarrow, bool kFull, int kL>
__global__ __launch_bounds__(kApplyThreads) void k_apply_tiled(
    uint64_t N, uint32_t ntiles, uint32_t tiles_per_xcd, uint32_t ngrp, uint32_t nchunk,
    uint32_t max_p, uint32_t zero_row, const ApplyGroup *__restrict__ grp, const ApplyChunk *__restrict__ chunks,
    const uint32_t *__restrict__ ug, const uint32_t *__restrict__ pzl,
    const uint32_t *__restrict__ ptptr, const uint64_t *__restrict__ ptz,
    const double2 *__restrict__ ptc, const void *__restrict__ xin, void *__restrict__ yout,
    double *__restrict__ partials) {
    using Shape = ApplyShape<kMixed, kL>;
    constexpr int L = Shape::L, R = Shape::R;
    // complex64 full tiles: paired rows (see gather); the patterns are built
    // in the paired-row bit order to match
    constexpr bool kPairs = kMixed && kFull;
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];

    const uint32_t tid = threadIdx.x;
    const uint32_t v = blockIdx.x / ntiles;
    uint32_t tile = blockIdx.x % ntiles;
    if (tiles_per_xcd) tile = (tile % 8) * tiles_per_xcd + tile / 8;
    const uint64_t tile_rows = kFull ? (uint64_t(1) << L) : (N < (uint64_t(1) << L) ? N : (uint64_t(1) << L));
    const uint64_t base = uint64_t(tile) * tile_rows;
    const cT *__restrict__ xv = reinterpret_cast<const cT *>(xin) + uint64_t(v) * N;

    size_t oA, oZ, oU;
    apply_lds_offsets<cT>(tile_rows, max_p, R, 0, &oA, &oZ, &oU);
    cT *xt = reinterpret_cast<cT *>(smem_raw);
    double2 *sA = reinterpret_cast<double2 *>(smem_raw + oA);
    uint32_t *sZ = reinterpret_cast<uint32_t *>(smem_raw + oZ);
    double2 *sU = reinterpret_cast<double2 *>(smem_raw + oU);

    for (uint64_t i = tid; i < tile_rows; i += kApplyThreads) xt[i] = xv[base + i];

    // accumulators: FP64 (ar, ai) or FP32 running sum s, compensation c, block partial b
    double ar[R], ai[R];
    float sr[R], si[R], br[R], bi[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        ar[k] = ai[k] = 0.0;
        sr[k] = si[k] = br[k] = bi[k] = 0.0f;
    }

    uint32_t since_flush = 0;  // groups in the FP32 block partials
    // add the FP32 block partials into the running sums with TwoSum
    auto flush = [&]() {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            float e, f;
            two_sum(sr[k], br[k], sr[k], e);
            two_sum(si[k], bi[k], si[k], f);
            br[k] = e;  // the rounding error starts the next block partial
            bi[k] = f;
        }
    };
    cT xg[R];
    for (uint32_t c = 0; c < nchunk; ++c) {
        const ApplyChunk ch = chunks[c];
        if (c) __syncthreads();  // the previous chunk's tables are no longer read
        chunk_tables<kNarrow, R>(ch, tid, base, zero_row, grp, ug, pzl, ptptr, ptz, ptc, sA, sZ, sU);
        // one loop per group class, each with a single straight path, in
        // blocks of at most kCompBlock groups flushed after each block
        auto run_class = [&](auto far_t, auto rowvar_t, auto cplx_t, auto swap_t, uint32_t g0, uint32_t g1) {
            constexpr bool kFar = decltype(far_t)::value, kRowvar = decltype(rowvar_t)::value;
            constexpr bool kCplx = decltype(cplx_t)::value, kSwap = decltype(swap_t)::value;
            for (uint32_t gb = g0; gb < g1;) {
                // up to the next flush point, counted over all classes
                const uint32_t room = uint32_t(kCompBlock) - since_flush;
                const uint32_t ge = g1 - gb > room ? gb + room : g1;
                for (uint32_t g = gb; g < ge; ++g) {
                    const ApplyGroup gd = grp[g];
                    // all descriptor fields in SGPRs before the gathers are issued,
                    // so that no later scalar-load wait also waits for them
                    asm volatile("" ::"s"(gd.p0), "s"(gd.ps), "s"(gd.pk), "s"(gd.p1), "s"(gd.u),
                                 "s"(gd.flags));
                    const auto fetch = [&]() {
                        gather<kFar, kFull, kPairs, R>(gd.m, tid, tile_rows, base, xt, xv, xg);
                    };
                    accumulate<kMixed, kFull, kRowvar, kCplx, kSwap, R>(gd, tid, tile_rows, zero_row, sA, sZ,
                                                                        sU, xg, fetch, ar, ai, br, bi);
                }
                since_flush += ge - gb;
                gb = ge;
                // reset in both precisions: room must stay >= 1 for the loop to advance
                if (since_flush == uint32_t(kCompBlock)) {
                    if (kMixed) flush();
                    since_flush = 0;
                }
            }
        };
        using F = std::false_type;
        using T = std::true_type;
        run_class(F{}, F{}, F{}, F{}, ch.cg[0], ch.cg[1]);
        run_class(T{}, F{}, F{}, F{}, ch.cg[1], ch.cg[2]);
        run_class(F{}, T{}, F{}, F{}, ch.cg[2], ch.cg[3]);
        run_class(T{}, T{}, F{}, F{}, ch.cg[3], ch.cg[4]);
        run_class(F{}, F{}, T{}, F{}, ch.cg[4], ch.cg[5]);
        run_class(T{}, F{}, T{}, F{}, ch.cg[5], ch.cg[6]);
        run_class(F{}, T{}, T{}, F{}, ch.cg[6], ch.cg[7]);
        run_class(T{}, T{}, T{}, F{}, ch.cg[7], ch.cg[8]);
        if constexpr (kPairs) {
            // masks with row bit 0: each row's partner is the other half of its pair
            run_class(F{}, F{}, F{}, T{}, ch.cg[8], ch.cg[9]);
            run_class(T{}, F{}, F{}, T{}, ch.cg[9], ch.cg[10]);
            run_class(F{}, T{}, F{}, T{}, ch.cg[10], ch.cg[11]);
            run_class(T{}, T{}, F{}, T{}, ch.cg[11], ch.cg[12]);
            run_class(F{}, F{}, T{}, T{}, ch.cg[12], ch.cg[13]);
            run_class(T{}, F{}, T{}, T{}, ch.cg[13], ch.cg[14]);
            run_class(F{}, T{}, T{}, T{}, ch.cg[14], ch.cg[15]);
            run_class(T{}, T{}, T{}, T{}, ch.cg[15], ch.cg[16]);
        }
    }

    if (kMixed && since_flush) flush();
    double er = 0.0, ei = 0.0;
    cT *__restrict__ yv = yout ? reinterpret_cast<cT *>(yout) + uint64_t(v) * N : nullptr;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint64_t l = kPairs ? ((uint64_t(tid) << 1) | uint64_t(k & 1) |
                                     (uint64_t(k >> 1) << (kApplyThreadBits + 1)))
                                  : tid + uint64_t(k) * kApplyThreads;
        if (!kFull && l >= tile_rows) break;
        double yr, yi;
        if (!kMixed) {
            yr = ar[k];
            yi = ai[k];
            if (yv) store_y(reinterpret_cast<double2 *>(yv) + base + l, make_double2(yr, yi));
        } else {
            yr = double(sr[k]) + double(br[k]);
            yi = double(si[k]) + double(bi[k]);
            if (yv) {
                if constexpr (kPairs) {
                    // the pair's two rows in one 16-byte store, at the odd row
                    if (k & 1) {
                        typedef float f4v __attribute__((ext_vector_type(4)));
                        const f4v pv = {sr[k - 1] + br[k - 1], si[k - 1] + bi[k - 1],
                                        sr[k] + br[k], si[k] + bi[k]};
                        __builtin_nontemporal_store(
                            pv, reinterpret_cast<f4v *>(reinterpret_cast<float2 *>(yv) + base + l - 1));
                    }
                } else {
                    store_y(reinterpret_cast<float2 *>(yv) + base + l,
                            make_float2(sr[k] + br[k], si[k] + bi[k]));
                }
            }
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

// Partner-staged form of k_apply_tiled (full tiles only).  Instead of one
// tile of x in LDS and far gathers from global memory, two tile buffers take
// the partner tiles t ^ hi in turn, one per run of groups sharing hi (hi = 0:
// the tile itself), so a partner tile is read once per chunk however many
// groups use it.  The load of run r + 1 (LDS-DMA: global memory straight into
// LDS, no registers) is in flight while run r is computed.

// One LDS-DMA load: each lane reads the 16 bytes at its `src`, and the wave
// writes them to LDS at lds + lane * 16 (lds: LDS byte address, wave-uniform,
// passed in M0; one wait state between the M0 write and the load).  Written
// as inline asm so that the compiler, which cannot tell the buffer being
// filled from the one being read, does not wait for the load before every
// later LDS read; the caller waits for it (vmcnt) and synchronises the block
// before reading what it wrote.  It has no register result.
// M0 is a reserved register that the compiler may hold a value in across the
// statement, so the asm saves it in an SGPR of its own and restores it after
// the load is issued (the load reads M0 at issue: back-to-back loads with
// different M0 values each land at their own address).
__device__ __forceinline__ void lds_dma16(const void *src, uint32_t lds) {
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %2\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %1, off\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(saved)
        : "v"(src), "s"(lds)
        : "memory");
}

typedef __attribute__((address_space(3))) unsigned char lds_byte;

template <bool kMixed, bool kNarrow, int kL>
__global__ __launch_bounds__(kApplyThreads) void k_apply_staged(
    uint64_t N, uint32_t ntiles, uint32_t tiles_per_xcd, uint32_t nchunk, uint32_t nrun,
    uint32_t max_p, uint32_t zero_row, const ApplyGroup *__restrict__ grp,
    const ApplyChunk *__restrict__ chunks, const ApplyRun *__restrict__ runs,
    const uint32_t *__restrict__ ug, const uint32_t *__restrict__ pzl,
    const uint32_t *__restrict__ ptptr, const uint64_t *__restrict__ ptz,
    const double2 *__restrict__ ptc, const void *__restrict__ xin, void *__restrict__ yout,
    double *__restrict__ partials) {
    using Shape = ApplyShape<kMixed, kL>;
    constexpr int L = Shape::L, R = Shape::R;
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    constexpr uint64_t tile_rows = uint64_t(1) << L;
    constexpr uint32_t kStageTileBytes = uint32_t(Shape::kTile);  // one stage buffer: one tile
    constexpr uint32_t kLoads = kStageTileBytes / (16u * kApplyThreads);  // 16-byte loads per thread
    static_assert(tile_rows * sizeof(cT) == kStageTileBytes, "a tile fills one stage buffer");
    static_assert(kStageTileBytes % (16u * kApplyThreads) == 0, "tile must split into 16-byte loads");
    __shared__ __align__(16) unsigned char bufs[2 * kStageTileBytes];
    extern __shared__ __align__(16) unsigned char smem_raw[];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = uint32_t(__builtin_amdgcn_readfirstlane(int(tid / 64)));
    const uint32_t v = blockIdx.x / ntiles;
    uint32_t tile = blockIdx.x % ntiles;
    if (tiles_per_xcd) tile = (tile % 8) * tiles_per_xcd + tile / 8;
    const uint64_t base = uint64_t(tile) * tile_rows;
    const cT *__restrict__ xv = reinterpret_cast<const cT *>(xin) + uint64_t(v) * N;

    size_t oA, oZ, oU;
    apply_lds_offsets<cT>(0, max_p, R, 0, &oA, &oZ, &oU);
    double2 *sA = reinterpret_cast<double2 *>(smem_raw + oA);
    uint32_t *sZ = reinterpret_cast<uint32_t *>(smem_raw + oZ);
    double2 *sU = reinterpret_cast<double2 *>(smem_raw + oU);

    // partner tile base ^ hi into buffer b: each wave fills 1 KB per load,
    // lane i the 16 bytes at i * 16
    const uint32_t lds_bufs = uint32_t(uintptr_t((lds_byte *)bufs));  // LDS address of bufs
    auto stage = [&](uint64_t hi, uint32_t b) {
        const unsigned char *src = reinterpret_cast<const unsigned char *>(xv + (base ^ hi));
#pragma unroll
        for (uint32_t j = 0; j < kLoads; ++j) {
            const uint32_t off = (j * kApplyThreads + wave * 64u) * 16u;
            lds_dma16(src + off + (tid % 64u) * 16u, lds_bufs + b * kStageTileBytes + off);
        }
    };

    double ar[R], ai[R];
    float sr[R], si[R], br[R], bi[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        ar[k] = ai[k] = 0.0;
        sr[k] = si[k] = br[k] = bi[k] = 0.0f;
    }
    uint32_t since_flush = 0;
    auto flush = [&]() {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            float e, f;
            two_sum(sr[k], br[k], sr[k], e);
            two_sum(si[k], bi[k], si[k], f);
            br[k] = e;  // the rounding error starts the next block partial
            bi[k] = f;
        }
    };
    cT xg[R];
    // the groups of one run, all gathering from the tile in `xt`
    auto run_groups = [&](const ApplyRun &run, const cT *xt) {
        auto run_class = [&](auto rowvar_t, auto cplx_t, uint32_t g0, uint32_t g1) {
            constexpr bool kRowvar = decltype(rowvar_t)::value;
            constexpr bool kCplx = decltype(cplx_t)::value;
            for (uint32_t gb = g0; gb < g1;) {
                const uint32_t room = uint32_t(kCompBlock) - since_flush;
                const uint32_t ge = g1 - gb > room ? gb + room : g1;
                for (uint32_t g = gb; g < ge; ++g) {
                    const ApplyGroup gd = grp[g];
                    asm volatile("" ::"s"(gd.p0), "s"(gd.ps), "s"(gd.pk), "s"(gd.p1), "s"(gd.u),
                                 "s"(gd.flags));
                    const auto fetch = [&]() { gather<false, true, false, R>(gd.m, tid, tile_rows, base, xt, xv, xg); };
                    accumulate<kMixed, true, kRowvar, kCplx, false, R>(gd, tid, tile_rows, zero_row, sA, sZ,
                                                                       sU, xg, fetch, ar, ai, br, bi);
                }
                since_flush += ge - gb;
                gb = ge;
                if (since_flush == uint32_t(kCompBlock)) {
                    if (kMixed) flush();
                    since_flush = 0;
                }
            }
        };
        using F = std::false_type;
        using T = std::true_type;
        run_class(F{}, F{}, run.g[0], run.g[1]);
        run_class(T{}, F{}, run.g[1], run.g[2]);
        run_class(F{}, T{}, run.g[2], run.g[3]);
        run_class(T{}, T{}, run.g[3], run.g[4]);
    };
    if (nrun) stage(runs[0].hi, 0);
    uint32_t r = 0;
    for (uint32_t c = 0; c < nchunk; ++c) {
        const ApplyChunk ch = chunks[c];
        if (c) __syncthreads();  // the previous chunk's tables are no longer read
        chunk_tables<kNarrow, R>(ch, tid, base, zero_row, grp, ug, pzl, ptptr, ptz, ptc, sA, sZ, sU);
        for (; r < ch.r1; ++r) {
            // run r's tile has landed (each wave waits for its own loads, the
            // barrier for all of them), and no wave still reads run r - 1's
            // buffer, which the next load overwrites
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            __syncthreads();
            if (r + 1 < nrun) stage(runs[r + 1].hi, (r + 1) & 1u);
            run_groups(runs[r], reinterpret_cast<const cT *>(bufs + (r & 1u) * kStageTileBytes));
        }
    }
    if (kMixed && since_flush) flush();

    // y, and <x|y> with x of the thread's rows read from global memory (the
    // buffers hold partner tiles)
    constexpr bool kFull = true;
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
            if (yv) store_y(reinterpret_cast<double2 *>(yv) + base + l, make_double2(yr, yi));
        } else {
            yr = double(sr[k]) + double(br[k]);
            yi = double(si[k]) + double(bi[k]);
            if (yv) store_y(reinterpret_cast<float2 *>(yv) + base + l, make_float2(sr[k] + br[k], si[k] + bi[k]));
        }
        if (partials) {
            const cT xr = xv[base + l];
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

template <bool kMixed, bool kNarrow, int kL>
static int launch_tiled(const qr_op *op, DeviceTables *t, int slot, uint32_t nvec, const void *d_x,
                        void *d_y, double *partials, uint32_t ntiles, hipStream_t s) {
    using Shape = ApplyShape<kMixed, kL>;
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    const uint64_t N = uint64_t(1) << op->nq;
    const uint64_t tile_rows = N < (uint64_t(1) << Shape::L) ? N : (uint64_t(1) << Shape::L);
    const DeviceTables::Patterns &pt = t->pat[slot];
    // the row-bit table holds max_u rows plus one zero row
    const size_t shmem = apply_lds_offsets<cT>(tile_rows, pt.max_p, Shape::R, pt.max_u + 1,
                                               nullptr, nullptr, nullptr);
    if (shmem > kApplyMaxLDS) return fail(QR_ERR_SIZE, "apply: LDS stage too large");
    const uint32_t tpx = (ntiles % 8 == 0) ? ntiles / 8 : 0;
    const bool full = tile_rows == (uint64_t(1) << Shape::L);
    auto kern = full ? k_apply_tiled<kMixed, kNarrow, true, kL> : k_apply_tiled<kMixed, kNarrow, false, kL>;
    QR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    const uint64_t nblocks = uint64_t(ntiles) * nvec;
    if (nblocks > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "apply grid too large");
    hipLaunchKernelGGL(kern, dim3(unsigned(nblocks)), dim3(kApplyThreads), shmem, s, N, ntiles,
                       tpx, pt.ngrp, pt.nchunk, pt.max_p, pt.max_u, pt.grp, pt.chunk, pt.ug, pt.pzl, pt.ptptr, pt.ptz,
                       pt.ptc, d_x, d_y, partials);
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

template <bool kMixed, bool kNarrow, int kL>
static int launch_staged(const qr_op *op, DeviceTables *t, int slot, uint32_t nvec, const void *d_x,
                         void *d_y, double *partials, uint32_t ntiles, hipStream_t s) {
    using Shape = ApplyShape<kMixed, kL>;
    using cT = typename std::conditional<kMixed, float2, double2>::type;
    const uint64_t N = uint64_t(1) << op->nq;
    const DeviceTables::Patterns &pt = t->pat[slot];
    // dynamic LDS holds the tables; the two tile buffers are static
    const size_t shmem = apply_lds_offsets<cT>(0, pt.max_p, Shape::R, pt.max_u + 1, nullptr,
                                               nullptr, nullptr);
    if (shmem + 2 * Shape::kTile + 1024 > kApplyMaxLDS)
        return fail(QR_ERR_SIZE, "apply: LDS stage too large");
    const uint32_t tpx = (ntiles % 8 == 0) ? ntiles / 8 : 0;
    auto kern = k_apply_staged<kMixed, kNarrow, kL>;
    QR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
    const uint64_t nblocks = uint64_t(ntiles) * nvec;
    if (nblocks > 0x7FFFFFFFull) return fail(QR_ERR_SIZE, "apply grid too large");
    hipLaunchKernelGGL(kern, dim3(unsigned(nblocks)), dim3(kApplyThreads), shmem, s, N, ntiles, tpx,
                       pt.nchunk, pt.nrun, pt.max_p, pt.max_u, pt.grp, pt.chunk, pt.run, pt.ug, pt.pzl,
                       pt.ptptr, pt.ptz, pt.ptc, d_x, d_y, partials);
    QR_HIP_CHECK(hipGetLastError());
    return QR_OK;
}

// Which form applies the operator.  The partner-staged kernel needs full
// tiles, and pays off when far groups share partner tiles: it loads a partner
// tile once per run instead of gathering it once per group, but the load of a
// one-group run is not hidden behind much work, and while the state is small
// enough to stay in the caches the per-group gathers are cheap.  Measured
// with the tiles above (F far groups on D distinct partner offsets; time of
// the staged form against the gather form):
//   mixed (2^11-row tiles):
//     24-qubit k-local, 100..3000 terms: F/D 1.8..4.1, 1.24..1.46x faster;
//     28-qubit k-local, 10000 terms: 1.40x faster;
//     Heisenberg ring, 22 and 24 qubits: F/D 1, 0.96..1.01x;
//     20-qubit random (8 MB state): F/D 1.05..1.24 -> 0.80..1.00x,
//     2.3 -> 1.06x, 4.0 -> 1.12x;
//   FP64 (2^10-row tiles): faster in every case measured -- k-local
//     1.29..1.86x, 28-qubit k-local 1.40x, Heisenberg ring 1.07..1.10x,
//     20-qubit random (F/D 1.0..2.3) 1.11..1.36x.
//   small complex64 tiles (2^10 rows; only launches of at most four tiles
//   per CU take them: at most 16 MiB of state on 256 CUs): the gather form, faster in
//   every case measured -- 20 qubits, one state, gather / staged: the
//   reference's H10 1.58 / 1.64 ms, random 2000 terms 1.19 / 1.45, random
//   300 terms 0.20 / 0.28, k-local 3000 terms 0.43 / 0.46, k-local 500
//   terms 0.13 / 0.15, Heisenberg ring 0.026 / 0.029.
// Hence: FP64 staged whenever a group reaches outside the tile; mixed on
// small tiles gathered; mixed on default tiles staged when F >= 1.5 D for
// states of at least 32 MiB per vector, F >= 2 D below that.  The
// environment variable QR_APPLY_FORM (1 gather, 2 staged) overrides the
// choice, for tests and measurement.
static bool prefer_staged(const qr_op *op, int L, size_t elem_bytes, bool small_tiles) {
    const char *e = getenv("QR_APPLY_FORM");
    if (e && atoi(e) == 1) return false;
    if (e && atoi(e) == 2) return true;
    if (small_tiles) return false;
    // gx ascends, so groups with equal bits above the tile are adjacent
    uint64_t far = 0, distinct = 0, prev = 0;
    for (uint64_t g = 0; g < op->G; ++g) {
        const uint64_t hi = op->gx[g] >> L;
        if (!hi) continue;
        ++far;
        if (distinct == 0 || hi != prev) ++distinct;
        prev = hi;
    }
    if (far == 0) return false;
    if (elem_bytes == 16) return true;
    const bool big = (uint64_t(1) << op->nq) * elem_bytes >= (uint64_t(32) << 20);
    return big ? 2 * far >= 3 * distinct : far >= 2 * distinct;
}

// One memory pool per device for the per-launch <x|y> workspaces.  Its
// release threshold keeps freed blocks in the pool, so that after the first
// launch taking a buffer costs no device allocation.
static std::mutex g_pool_mu;
static hipMemPool_t g_pool[kMaxDevices];

static int partials_pool(hipMemPool_t *out) {
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (!g_pool[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t p;
        QR_HIP_CHECK(hipMemPoolCreate(&p, &props));
        uint64_t keep = ~uint64_t(0);
        QR_HIP_CHECK(hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep));
        g_pool[dev] = p;
    }
    *out = g_pool[dev];
    return QR_OK;
}

// Workspace of the allocation-free entry point (qr_pauli_apply_noalloc): the
// <x|y> block partials, 2 doubles per (tile, vector), in a file-scope device
// array of 4 MiB that each device holds until the process ends.  Every call
// on one device shares it and it keeps the last call's contents; each block
// overwrites its own two slots, so nothing needs clearing.  In host code the
// array's name is no device address: its address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kScratchDoubles = size_t(1) << 19;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices];

static int scratch_partials(double **out) {
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        QR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(QR_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return QR_OK;
}

// Compute units of the current device (queried once per device).
static int apply_cus(int *out) {
    static std::mutex mu;
    static int cache[kMaxDevices];
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!cache[dev]) {
        int cus = 0;
        QR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cache[dev] = cus > 0 ? cus : 1;
    }
    *out = cache[dev];
    return QR_OK;
}

// One way of applying an operator: tile size, form, and the pattern-table
// slot that holds its tables (2 * kind + staged; kind 0: FP64, 1: complex64,
// 2: small complex64 tiles).
struct ApplyPlan {
    int L = 0, slot = 0;
    bool staged = false, pairs = false, small = false;
    uint32_t pcap = 0, ucap = 0;
};

template <bool kMixed, int kL>
static ApplyPlan make_plan(const qr_op *op, int kind) {
    using Shape = ApplyShape<kMixed, kL>;
    ApplyPlan p;
    p.L = int(op->nq) < kL ? int(op->nq) : kL;
    p.staged = p.L == kL && prefer_staged(op, p.L, kMixed ? 8 : 16, kind == 2);
    p.pcap = p.staged ? Shape::kPatChunkS : Shape::kPatChunk;
    p.ucap = p.staged ? Shape::kTableRowsS : Shape::kTableRows;
    // the gather form on full complex64 tiles runs the paired-row layout
    p.pairs = kMixed && !p.staged && p.L == kL;
    p.small = kind == 2;
    p.slot = 2 * kind + (p.staged ? 1 : 0);
    return p;
}

int launch_pauli_apply(const qr_op *op, DeviceTables *t, int32_t precision, uint32_t nvec,
                       const void *d_x, void *d_y, double *d_expect, hipStream_t s,
                       bool noalloc) {
    // QR_APPLY_WIDE=1 runs the 64-bit-mask instances on any operator, so that
    // the tests can check them without a state of more than 32 qubits
    const char *wide_env = getenv("QR_APPLY_WIDE");
    const bool force_wide = wide_env && atoi(wide_env) == 1;
    const bool mixed = precision == QR_PREC_MIXED, narrow = op->nq <= 32 && !force_wide;
    const uint64_t N = uint64_t(1) << op->nq;
    if (nvec == 0) return fail(QR_ERR_SIZE, "apply: nvec == 0");
    int rc;
    ApplyPlan plan;
    if (!mixed) {
        plan = make_plan<false, ApplyShape<false>::L>(op, 0);
    } else {
        // small complex64 tiles when the default ones give fewer than four
        // blocks per CU (the blocks one CU holds at once); QR_APPLY_SMALL
        // (0 never, 1 always where the state has more rows than a small
        // tile) overrides the choice, for tests and measurement
        int cus = 0;
        if ((rc = apply_cus(&cus))) return rc;
        const bool can_small = int(op->nq) > kMixedSmallL;
        bool small = can_small && (N >> ApplyShape<true>::L) * nvec < uint64_t(4) * uint64_t(cus);
        const char *small_env = getenv("QR_APPLY_SMALL");
        if (small_env) small = can_small && atoi(small_env) == 1;
        plan = small ? make_plan<true, kMixedSmallL>(op, 2) : make_plan<true, ApplyShape<true>::L>(op, 1);
        // without allocation only uploaded tables can serve: those of the
        // other tile size when this one's are not on the device
        if (noalloc && can_small) {
            const ApplyPlan other = small ? make_plan<true, ApplyShape<true>::L>(op, 1)
                                          : make_plan<true, kMixedSmallL>(op, 2);
            if (!tables_ready(op, t, plan.slot, plan.L) && tables_ready(op, t, other.slot, other.L))
                plan = other;
        }
    }
    rc = device_patterns(op, t, plan.slot, plan.L, kApplyThreadBits, plan.pcap, plan.ucap,
                         plan.staged, plan.pairs, noalloc);
    if (rc) return rc;
    const int L = plan.L;
    const bool staged = plan.staged;
    if ((N >> L) * uint64_t(nvec) > 0x7FFFFFFFull)
        return fail(QR_ERR_SIZE, "apply: state too large for one launch");
    const uint32_t ntiles = uint32_t(N >> L);
    // <x|y> block partials: a buffer of this launch's own, taken from the
    // device's pool and given back after the reduction, both in stream order,
    // so launches on one operator from several streams never share one
    double *partials = nullptr;
    if (d_expect && noalloc) {
        const uint64_t need = uint64_t(ntiles) * nvec * 2;
        if (need > kScratchDoubles)
            return fail(QR_ERR_SIZE, "apply_noalloc: <x|H|x> of " + std::to_string(nvec) +
                                         " vectors of " + std::to_string(ntiles) +
                                         " tiles needs " + std::to_string(need * 8) +
                                         " bytes of partials; the scratch array holds " +
                                         std::to_string(kScratchDoubles * 8));
        if ((rc = scratch_partials(&partials))) return rc;
    } else if (d_expect) {
        hipMemPool_t pool;
        if ((rc = partials_pool(&pool))) return rc;
        QR_HIP_CHECK(hipMallocFromPoolAsync(reinterpret_cast<void **>(&partials),
                                            size_t(ntiles) * nvec * 2 * sizeof(double), pool, s));
    }
    constexpr int kLm = ApplyShape<true>::L, kLd = ApplyShape<false>::L, kLs = kMixedSmallL;
    const int sl = plan.slot;
    if (staged) {
        if (mixed && plan.small)
            rc = narrow ? launch_staged<true, true, kLs>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s)
                        : launch_staged<true, false, kLs>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s);
        else if (mixed)
            rc = narrow ? launch_staged<true, true, kLm>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s)
                        : launch_staged<true, false, kLm>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s);
        else
            rc = narrow ? launch_staged<false, true, kLd>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s)
                        : launch_staged<false, false, kLd>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s);
    } else if (mixed && plan.small) {
        rc = narrow ? launch_tiled<true, true, kLs>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s)
                    : launch_tiled<true, false, kLs>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s);
    } else if (mixed) {
        rc = narrow ? launch_tiled<true, true, kLm>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s)
                    : launch_tiled<true, false, kLm>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s);
    } else {
        rc = narrow ? launch_tiled<false, true, kLd>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s)
                    : launch_tiled<false, false, kLd>(op, t, sl, nvec, d_x, d_y, partials, ntiles, s);
    }
    if (!rc && d_expect) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(nvec), dim3(256), 0, s, ntiles, partials,
                           d_expect);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_fail(e, "k_reduce_partials");
    }
    if (partials && !noalloc) {
        const hipError_t e = hipFreeAsync(partials, s);
        if (!rc && e != hipSuccess) rc = hip_fail(e, "hipFreeAsync(partials)");
    }
    if (rc) return rc;
    return QR_OK;
}

}  // namespace qr
