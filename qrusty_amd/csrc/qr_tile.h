// qr_tile.h -- the LDS-staged row-tile builder shared by the single-operator
// CSR build (qr_csr.hip) and the batched build+matvec (qr_batch.hip).
//
// Replaces accel::rowwise::make_row (qrusty/src/accel.rs:171-210) for one tile
// of rows.  Thread layout: 1024 threads = RS rows x JW groups (JW a power of
// two >= min(G, 1024)); a tile is 16 batches of RS consecutive rows.  Thread
// (ro, lane) owns groups j = c*JW + lane (c < nchunks) and rows
// r = tile_base + m*RS + ro, m = 0..15.  All bits of r but the four m bits are
// fixed for the thread, so each group's rank list (qr_api.cpp:
// build_rank_lists) is walked once per tile: entries on fixed bits are summed,
// the at most four entries on m bits form an offset table.  Each batch is
// written into LDS at (row, rank) and copied out by consecutive threads taking
// consecutive entries, so every global store is fully coalesced.  Values are
// the term-order fold of (-1)^popcount(r & z_t) c'_t (first value, then +),
// bit-identical to the reference.
//
// With kMatvec the tile also produces y[r] = sum_g value * x[r ^ x_g] for its
// rows from the values still in registers (FP64 FMAs per thread, then a fixed
// reduction tree over the JW lanes of a row), i.e. a CSR matvec fused into the
// build that never reads the CSR back; without CSR outputs nothing is staged.
#pragma once

#include "qr_internal.h"
#include "qr_mv.h"

#include <type_traits>

namespace qr {

constexpr int kBuildM = 16;            // batches per tile (4 varying row bits)
constexpr int kStageThreads = 1024;
constexpr uint32_t kStageMaxG = 4096;  // 24 B per entry: 96 KB of LDS for one row
constexpr int kStageMaxChunks = 4;
// "lean" class: up to 6 chunks (G <= 6144, a 144 KB row stage); only the rank
// data and term offsets stay in registers, the first term of each group is
// re-read (from L2) for every batch
constexpr int kStageLeanChunks = 6;
constexpr uint32_t kStageMaxGLean = 6144;

template <bool kNarrow>
__device__ __forceinline__ uint32_t parity_and(uint64_t a, uint64_t b) {
    if (kNarrow) return __popc(uint32_t(a) & uint32_t(b)) & 1u;
    return __popcll(a & b) & 1u;
}

// (-1)^s * c, by flipping sign bits (exact, like the reference's -coeff)
__device__ __forceinline__ double2 signed_coeff(double2 c, uint32_t s) {
    if (s) { c.x = -c.x; c.y = -c.y; }
    return c;
}

// One operator's view for the tile builder.
struct TileOp {
    const uint64_t *gx;
    const uint32_t *gterm;
    const uint64_t *tz;
    const double2 *tc;
    const uint32_t *rl_ptr;
    const RankEntry *rl;
    uint64_t *indices;   // CSR outputs at (r - row_begin) * G; may be null
    double2 *data;
    uint64_t *indptr;    // indptr[r - row_begin] = r * G; may be null
    const double2 *x;    // fused matvec input / output (kMatvec only)
    double2 *y;
    uint64_t row_begin, row_end;
    uint32_t G, jw_log2, nchunks, dbuf;
    const uint32_t *heavy;  // groups folded once per row into LDS (ascending)
    uint32_t nheavy;        // 0: none
    // Sub-operator of a split operator: the groups [gbase, gbase + G) of a
    // subtree of the mask trie.  In row r they occupy the contiguous ranks
    // [off(r), off(r) + G), off(r) = gbase + the sigma of the subtree's
    // nanc ancestors whose bit is set in r, i.e. the first nanc entries of
    // its first group's rank list.  gx, gterm and rl_ptr point at group
    // gbase; rows are gstride entries apart in the outputs.  An operator
    // built whole has gbase = nanc = 0 and gstride = G.
    uint32_t gbase, nanc, gstride;
    // tail table (qr_op::ttz, ttc), pointing at group gbase: term 2 + u of
    // group j at u * gstride + j, for u < ntail; ntail = 0: none
    const uint64_t *ttz;
    const double2 *ttc;
    uint32_t ntail;
    // heavy-group instances: each block builds 2^fold_log2 consecutive tiles
    // and folds the heavy groups once for all of their rows
    uint32_t fold_log2;
};

// The arrays of a TileOp seen through pointers typed as global memory.  A
// TileOp that reaches the kernel inside a descriptor in memory (the batch)
// holds generic pointers, which the compiler would access with flat
// instructions: 64-bit per-lane addresses, counted against both vmcnt and
// lgkmcnt.  The builder reads and writes only through this view.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
struct TileMem {
    gptr<const uint64_t> gx;
    gptr<const uint32_t> gterm;
    gptr<const uint64_t> tz;
    gptr<const dvec2> tc;
    gptr<const uint32_t> rl_ptr;
    gptr<const RankEntry> rl;
    gptr<uint64_t> indices;
    gptr<dvec2> data;
    gptr<uint64_t> indptr;
    gptr<const dvec2> x;
    gptr<dvec2> y;
    gptr<const uint32_t> heavy;
    gptr<const uint64_t> ttz;
    gptr<const dvec2> ttc;
};
__device__ __forceinline__ TileMem tile_mem(const TileOp &op) {
    return {as_global(op.gx),
            as_global(op.gterm),
            as_global(op.tz),
            as_global(reinterpret_cast<const dvec2 *>(op.tc)),
            as_global(op.rl_ptr),
            as_global(op.rl),
            as_global(op.indices),
            as_global(reinterpret_cast<dvec2 *>(op.data)),
            as_global(op.indptr),
            as_global(reinterpret_cast<const dvec2 *>(op.x)),
            as_global(reinterpret_cast<dvec2 *>(op.y)),
            as_global(op.heavy),
            as_global(op.ttz),
            as_global(reinterpret_cast<const dvec2 *>(op.ttc))};
}
__device__ __forceinline__ RankEntry ld_rank(gptr<const RankEntry> p) {
    RankEntry e;
    e.bit = p->bit;
    e.sigma = p->sigma;
    return e;
}
__device__ __forceinline__ double2 ld2(gptr<const dvec2> p) {
    const dvec2 v = *p;
    return make_double2(v.x, v.y);
}

// LDS bytes for one operator's stage: (one or two) x (index + value) x batch,
// plus 2 x 16 wave partials for the fused matvec.  The stage can be
// double-buffered (the copy-out of batch m overlaps the build of m+1) when both
// buffers fit in QR_STAGE_DBUF_MAX bytes; by default it is not: a single stage
// (48 KB at G = 2000) lets two 1024-thread blocks share a CU, which overlaps
// one block's copy-out with the other's build and measured 20% faster.
#ifndef QR_STAGE_DBUF_MAX
#define QR_STAGE_DBUF_MAX 0
#endif
__host__ __device__ inline size_t stage_bytes_for(uint32_t G, uint32_t jw_log2, uint32_t dbuf) {
    const size_t batch_pad = ((size_t(G) << (10 - jw_log2)) + 1) & ~size_t(1);
    const size_t one = batch_pad * (sizeof(uint64_t) + sizeof(double2));
    return (dbuf ? 2 * one : one) + 2 * 16 * sizeof(double2);
}
__host__ __device__ inline size_t stage_bytes(uint32_t G, uint32_t jw_log2, uint32_t *dbuf_out) {
    const uint32_t dbuf = stage_bytes_for(G, jw_log2, 1) - 2 * 16 * sizeof(double2) <= QR_STAGE_DBUF_MAX ? 1u : 0u;
    if (dbuf_out) *dbuf_out = dbuf;
    return stage_bytes_for(G, jw_log2, dbuf);
}

// LDS bytes of the heavy-group values: one double2 per row of the tile and
// heavy group, placed after the stage.  At most kHeavyMaxBytes; the launch
// takes as many heavy groups as fit.
constexpr size_t kHeavyMaxBytes = 32 * 1024;
__host__ __device__ inline size_t heavy_bytes(uint32_t nheavy, uint32_t jw_log2) {
    return size_t(nheavy) * (size_t(kBuildM) << (10 - jw_log2)) * sizeof(double2);
}
__host__ __device__ inline uint32_t heavy_fit(uint32_t nheavy, uint32_t jw_log2) {
    while (nheavy && heavy_bytes(nheavy, jw_log2) > kHeavyMaxBytes) --nheavy;
    return nheavy;
}
// Tiles per block (log2, at most kFoldMaxLog2) of a heavy-group instance that
// builds several (fold_tiles), as many as keep the stage and their heavy
// values within kFoldBudget bytes of LDS (those instances run one block per
// CU for their registers) and leave at least kFoldMinBlocks blocks of the
// launch's ntiles tiles, so that small operators still fill the chip.
// up to 8 tiles: H12 (2^19 rows) 4.66 -> 4.42 ms against up to 4, H10 and
// H11 unchanged; 16 do not fit the LDS budget there
#ifndef QR_FOLD_MAX_LOG2
#define QR_FOLD_MAX_LOG2 3
#endif
constexpr size_t kFoldBudget = 150 * 1024;
// The two-chunk whole-operator instances (one block per CU, so no second
// block overlaps a batch's copy-out with the next batch's build)
// double-buffer the stage when it and one tile's heavy values fit the
// budget; the tiles per block are then chosen with the double stage.
// Measured on H12 (2^19 rows): 3.93 -> 3.72-3.76 ms.
#ifndef QR_HEAVY_DBUF
#define QR_HEAVY_DBUF 1
#endif
__host__ __device__ inline size_t heavy_stage(bool several, uint32_t G, uint32_t jw_log2, size_t heavy,
                                              uint32_t *dbuf) {
    *dbuf = 0;
    if (QR_HEAVY_DBUF && several && heavy && stage_bytes_for(G, jw_log2, 1) + heavy <= kFoldBudget) *dbuf = 1;
    return stage_bytes_for(G, jw_log2, *dbuf);
}
constexpr uint32_t kFoldMaxLog2 = QR_FOLD_MAX_LOG2;
constexpr uint64_t kFoldMinBlocks = 512;
__host__ __device__ inline uint32_t fold_tiles_log2(bool several, size_t stage, size_t heavy, uint64_t ntiles) {
    if (!several || heavy == 0) return 0;
    uint32_t f = kFoldMaxLog2;
    while (f && (stage + (heavy << f) > kFoldBudget || (ntiles >> f) < kFoldMinBlocks)) --f;
    return f;
}

// Term-order fold of group j in row r: (-1)^popcount(r & z_t) c'_t, the
// first value, then + (accel.rs:191-208).  Terms are loaded kF at a time,
// all kF loads issued before the first add; the adds stay in term order.
// Four at a time measured 2.68 -> 2.58 ms on the 24-qubit Heisenberg build;
// eight cost that (one-chunk) instance its second block per CU, so eight are
// used only by the two-chunk instances, which run one block per CU anyway.
template <bool kNarrow, uint32_t kF>
__device__ __forceinline__ double2 fold_group(const TileMem &mem, uint32_t j, uint64_t r) {
    const uint32_t t0 = mem.gterm[j], t1 = mem.gterm[j + 1];
    double2 v = signed_coeff(ld2(mem.tc + t0), parity_and<kNarrow>(r, mem.tz[t0]));
    uint32_t t = t0 + 1;
    for (; t + kF <= t1; t += kF) {
        double2 ct[kF];
#pragma unroll
        for (uint32_t u = 0; u < kF; ++u) ct[u] = signed_coeff(ld2(mem.tc + t + u), parity_and<kNarrow>(r, mem.tz[t + u]));
#pragma unroll
        for (uint32_t u = 0; u < kF; ++u) {
            v.x = v.x + ct[u].x;
            v.y = v.y + ct[u].y;
        }
    }
    for (; t < t1; ++t) {
        const double2 ct = signed_coeff(ld2(mem.tc + t), parity_and<kNarrow>(r, mem.tz[t]));
        v.x = v.x + ct.x;
        v.y = v.y + ct.y;
    }
    return v;
}

// The same fold for NR rows of one group at once: each term is loaded once
// for all of them, and the NR sums are independent chains (each in term
// order, so every value equals fold_group's).
template <bool kNarrow, uint32_t kF, int NR>
__device__ __forceinline__ void fold_group_rows(const TileMem &mem, uint32_t j, const uint64_t (&r)[NR],
                                                double2 (&v)[NR]) {
    const uint32_t t0 = mem.gterm[j], t1 = mem.gterm[j + 1];
    {
        const double2 c = ld2(mem.tc + t0);
        const uint64_t z = mem.tz[t0];
#pragma unroll
        for (int q = 0; q < NR; ++q) v[q] = signed_coeff(c, parity_and<kNarrow>(r[q], z));
    }
    uint32_t t = t0 + 1;
    for (; t + kF <= t1; t += kF) {
        double2 ct[kF];
        uint64_t zt[kF];
#pragma unroll
        for (uint32_t u = 0; u < kF; ++u) {
            ct[u] = ld2(mem.tc + t + u);
            zt[u] = mem.tz[t + u];
        }
#pragma unroll
        for (uint32_t u = 0; u < kF; ++u) {
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                const double2 e = signed_coeff(ct[u], parity_and<kNarrow>(r[q], zt[u]));
                v[q].x = v[q].x + e.x;
                v[q].y = v[q].y + e.y;
            }
        }
    }
    for (; t < t1; ++t) {
        const double2 c = ld2(mem.tc + t);
        const uint64_t z = mem.tz[t];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const double2 e = signed_coeff(c, parity_and<kNarrow>(r[q], z));
            v[q].x = v[q].x + e.x;
            v[q].y = v[q].y + e.y;
        }
    }
}

__host__ __device__ inline uint32_t stage_jw_log2(uint32_t G) {
    uint32_t jw = 0;
    while ((1u << jw) < G && jw < 10) ++jw;
    return jw;
}

// kMode: kTilePlain; kTileHeavy, the operator has heavy groups
// (op.nheavy > 0); kTileSub, a sub-operator (op.nanc, op.gbase, op.gstride);
// kTileSubHeavy, a sub-operator with heavy groups (leaf-local indices).
// Separate instantiations, so that plain operators run the plain loop (its
// register count decides whether two blocks share a CU).
constexpr int kTilePlain = 0, kTileHeavy = 1, kTileSub = 2, kTileSubHeavy = 3;
// Heavy-group instances that build several tiles per block, folding the heavy
// groups once for all of them: the two-chunk ones (one block per CU anyway)
// and the one-chunk sub-operators, which the tile loop takes from 60 to 78
// VGPRs and one block per CU.  Measured (ms): H10 16.75 -> 16.45 and H11
// 13.92 -> 13.48 with the one-chunk sub-operators in; the whole one-chunk
// operators (the 24-qubit Heisenberg build) stay at one tile per block
// (2.25 ms against 3.03 with it).
template <int kChunks, int kMode>
__host__ __device__ constexpr bool fold_tiles() {
    return (kMode == kTileHeavy && kChunks >= 2) || kMode == kTileSubHeavy;
}
// `sub` is the tile's place among the 2^op.fold_log2 tiles of its block
// (heavy-group instances; 0 otherwise): the first folds the heavy groups for
// all of them.
template <bool kNarrow, int kChunks, bool kMatvec, int kMode = kTilePlain>
__device__ __forceinline__ void build_tile_staged(const TileOp &op, uint64_t tile,
                                                  unsigned char *smem_raw, uint32_t sub = 0) {
    constexpr bool kHeavy = kMode == kTileHeavy || kMode == kTileSubHeavy;
    constexpr bool kSub = kMode == kTileSub || kMode == kTileSubHeavy;
    // the second term of each group stays in registers when few chunks leave
    // room; with more than kStageMaxChunks not even the first one does.  Not
    // in the sub-operator mode without heavy groups (random operators of more
    // than 2048 groups, and every operator of the batch's split form), whose
    // groups rarely have a second term: there the held term cost the batch's
    // two-chunk kernel 10 of its 14 spilled registers (256-operator batch
    // 79.6-79.9 -> 78.0-78.3 ms)
    constexpr bool kHold2 = kChunks <= 2 && kMode != kTileSub;
    // instances of operators with heavy groups (whose other groups run to 4-6
    // terms) that run one block per CU also hold terms 2 and 3, loaded after
    // the heavy fold so that they do not add to its registers (narrow masks:
    // the 64-bit ones would spill)
    constexpr bool kHold4 = kHold2 && kNarrow && fold_tiles<kChunks, kMode>();
    // the one-chunk sub-operators have room for terms 4 and 5 as well
    constexpr bool kHold6 = kHold4 && kChunks == 1;
    constexpr bool kLean = kChunks > kStageMaxChunks;
    const TileMem mem = tile_mem(op);
    const uint32_t G = op.G, jw_log2 = op.jw_log2, nchunks = op.nchunks;
    const uint64_t row_begin = op.row_begin, row_end = op.row_end;
    const uint32_t rs_log2 = 10u - jw_log2;
    const uint32_t tile_log2 = rs_log2 + 4u;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & ((1u << jw_log2) - 1u);
    const uint32_t ro = tid >> jw_log2;
    const uint64_t tile_base = ((row_begin >> tile_log2) + tile) << tile_log2;
    const uint64_t rfix = tile_base | ro;
    const uint32_t batch = G << rs_log2;  // entries per batch of RS rows
    const uint32_t batch_pad = (batch + 1u) & ~1u;
    // stage b (index + value of one batch) at byte b * stage_bytes; the
    // pointers are formed from smem_raw and a byte offset (not picked from an
    // array of pointers), so that the compiler keeps them in the LDS address
    // space: through generic pointers the stage accesses become flat
    // instructions, which count against vmcnt, and every copy-out load then
    // waited for all earlier global stores
    const uint32_t stage_b = batch_pad * 24u;
    double2 *red = reinterpret_cast<double2 *>(smem_raw + (op.dbuf ? 2u : 1u) * stage_b);
    // heavy-group values of the rows of the block's tiles:
    // hv[h << (tile_log2 + fold_log2) | sub << tile_log2 | (r - tile_base)]
    double2 *hv = reinterpret_cast<double2 *>(smem_raw + stage_bytes_for(G, jw_log2, op.dbuf));
    const uint32_t fold_log2 = fold_tiles<kChunks, kMode>() ? op.fold_log2 : 0u;

    // rows of the tile: indptr
    if (op.indptr) {
        const uint64_t tr = uint64_t(1) << tile_log2;
        const uint64_t stride = kSub ? op.gstride : G;
        for (uint64_t i = tid; i < tr; i += kStageThreads) {
            const uint64_t r = tile_base + i;
            if (r >= row_begin && r < row_end) mem.indptr[r - row_begin] = r * stride;
            if (r + 1 == row_end) mem.indptr[r + 1 - row_begin] = (r + 1) * stride;
        }
    }
    // a sub-operator's first rank in the row: fixed bits summed, batch bits
    // as four offsets (like the rank lists below).  The same for every
    // thread of the block -- a sub-operator has one row per batch (JW = 1024),
    // a whole operator no ancestors -- so it is held in scalar registers.
    int32_t off_fix = 0, o0 = 0, o1 = 0, o2 = 0, o3 = 0;
    if (kSub) off_fix = int32_t(op.gbase);
    for (uint32_t e = 0; kSub && e < op.nanc; ++e) {
        const RankEntry en = ld_rank(mem.rl + mem.rl_ptr[0] + e);
        const uint32_t b = uint32_t(en.bit);
        const uint32_t q = b - rs_log2;  // wraps for b < rs_log2
        if (q == 0) o0 = en.sigma;
        else if (q == 1) o1 = en.sigma;
        else if (q == 2) o2 = en.sigma;
        else if (q == 3) o3 = en.sigma;
        else if ((rfix >> b) & 1) off_fix += en.sigma;
    }
    if (kSub) {
        off_fix = __builtin_amdgcn_readfirstlane(off_fix);
        o0 = __builtin_amdgcn_readfirstlane(o0);
        o1 = __builtin_amdgcn_readfirstlane(o1);
        o2 = __builtin_amdgcn_readfirstlane(o2);
        o3 = __builtin_amdgcn_readfirstlane(o3);
    }

    // masks of <= 32-qubit operators are held in 32 bits
    using mask_t = typename std::conditional<kNarrow, uint32_t, uint64_t>::type;
    // rank offsets of the four batch bits, two 16-bit values per register
    // (|sigma| < G <= kStageMaxGLean)
    int32_t rfx[kChunks], s01[kChunks], s23[kChunks];
    mask_t xj[kChunks], z0[kChunks], z1[kChunks];
    double2 c0[kChunks], c1[kChunks];
    uint32_t t0[kChunks], t1[kChunks];
    int32_t hs[kChunks];  // heavy slot of group j, or -1
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const uint32_t j = (uint32_t(c) << jw_log2) + lane;
        rfx[c] = int32_t(kSub ? op.gbase + j : j);
        hs[c] = -1;
        if (kHeavy) {
            // heavy groups are listed ascending: binary search
            uint32_t lo = 0, hi = op.nheavy;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (mem.heavy[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            if (lo < op.nheavy && mem.heavy[lo] == j) hs[c] = int32_t(lo);
        }
        int32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        s01[c] = s23[c] = 0;
        xj[c] = z0[c] = z1[c] = 0;
        c0[c] = c1[c] = make_double2(0.0, 0.0);
        t0[c] = t1[c] = 0;
        if (uint32_t(c) >= nchunks || j >= G) continue;
        const uint32_t e1 = mem.rl_ptr[j + 1];
        for (uint32_t e = mem.rl_ptr[j]; e < e1; ++e) {
            const RankEntry en = ld_rank(mem.rl + e);
            const uint32_t b = uint32_t(en.bit);
            const uint32_t q = b - rs_log2;  // wraps for b < rs_log2
            if (q == 0) s0 = en.sigma;
            else if (q == 1) s1 = en.sigma;
            else if (q == 2) s2 = en.sigma;
            else if (q == 3) s3 = en.sigma;
            else if ((rfix >> b) & 1) rfx[c] += en.sigma;
        }
        s01[c] = int32_t((uint32_t(s1) << 16) | (uint32_t(s0) & 0xffffu));
        s23[c] = int32_t((uint32_t(s3) << 16) | (uint32_t(s2) & 0xffffu));
        xj[c] = mask_t(mem.gx[j]);
        t0[c] = mem.gterm[j];
        t1[c] = mem.gterm[j + 1];
        if (!kLean) {
            z0[c] = mask_t(mem.tz[t0[c]]);
            c0[c] = ld2(mem.tc + t0[c]);
        }
        if (kHold2 && t1[c] > t0[c] + 1) {
            z1[c] = mask_t(mem.tz[t0[c] + 1]);
            c1[c] = ld2(mem.tc + t0[c] + 1);
        }
    }

    if (kHeavy && sub == 0) {
        // heavy groups: one thread per (group, row) folds the whole group, so
        // no lane of a batch waits on a long term chain; the block's first
        // tile folds for the rows of all its tiles, so that the longest
        // chain, which sets the time of the pass, is paid once for them
        const uint32_t tr = uint32_t(kBuildM) << (rs_log2 + fold_log2);
        if constexpr (fold_tiles<kChunks, kMode>()) {
            // (one block per CU) two rows of a group per thread, tr/2 apart:
            // one load of each term for both, two independent add chains
            constexpr int NR = 2;
            const uint32_t half = tr / NR;
            for (uint32_t i = tid; i < op.nheavy * half; i += kStageThreads) {
                const uint32_t g = i / half, o = i % half;
                uint64_t rr[NR];
                double2 vv[NR];
#pragma unroll
                for (int q = 0; q < NR; ++q) rr[q] = tile_base + o + uint64_t(q) * half;
                fold_group_rows<kNarrow, 4u, NR>(mem, mem.heavy[g], rr, vv);
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (rr[q] >= row_begin && rr[q] < row_end) hv[g * tr + o + uint32_t(q) * half] = vv[q];
            }
        } else {
            for (uint32_t i = tid; i < op.nheavy * tr; i += kStageThreads) {
                const uint64_t r = tile_base + (i & (tr - 1));
                if (r >= row_begin && r < row_end)
                    hv[i] = fold_group<kNarrow, kChunks >= 2 ? 8u : 4u>(mem, mem.heavy[i / tr], r);
            }
        }
        __syncthreads();
    }
    mask_t z2[kChunks], z3[kChunks], z4[kChunks], z5[kChunks];
    double2 c2[kChunks], c3[kChunks], c4[kChunks], c5[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        z2[c] = z3[c] = z4[c] = z5[c] = 0;
        c2[c] = c3[c] = c4[c] = c5[c] = make_double2(0.0, 0.0);
        const uint32_t j = (uint32_t(c) << jw_log2) + lane;
        if (!kHold4 || uint32_t(c) >= nchunks || j >= G) continue;
        // term 2 + u from the tail table (u < ntail) or the group's own range
        auto hold = [&](uint32_t u, mask_t &z, double2 &cc) {
            if (t1[c] <= t0[c] + 2 + u) return;
            if (u < op.ntail) {
                z = mask_t(mem.ttz[u * op.gstride + j]);
                cc = ld2(mem.ttc + u * op.gstride + j);
            } else {
                z = mask_t(mem.tz[t0[c] + 2 + u]);
                cc = ld2(mem.tc + t0[c] + 2 + u);
            }
        };
        hold(0, z2[c], c2[c]);
        hold(1, z3[c], c3[c]);
        if (kHold6) {
            hold(2, z4[c], c4[c]);
            hold(3, z5[c], c5[c]);
        }
    }

#pragma unroll 1
    for (int m = 0; m < kBuildM; ++m) {
        const int buf = op.dbuf ? (m & 1) : 0;
        uint64_t *si = reinterpret_cast<uint64_t *>(smem_raw + uint32_t(buf) * stage_b);
        double2 *sd = reinterpret_cast<double2 *>(si + batch_pad);
        const uint64_t r = rfix + (uint64_t(m) << rs_log2);
        const bool row_ok = r >= row_begin && r < row_end;
        const int32_t off = kSub ? off_fix + ((m & 1) ? o0 : 0) + ((m & 2) ? o1 : 0) +
                                       ((m & 4) ? o2 : 0) + ((m & 8) ? o3 : 0)
                                 : 0;
        // fused matvec: this thread's entries of row r, from registers.  Its
        // groups are the same in every batch, so across the tile's batches it
        // gathers x[r ^ x_j] for rows r that differ in the batch bits only:
        // the same few cache lines.
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const uint32_t j = (uint32_t(c) << jw_log2) + lane;
            if (uint32_t(c) >= nchunks || j >= G || !row_ok) continue;
            const int32_t rank = rfx[c] + ((m & 1) ? int32_t(int16_t(s01[c])) : 0) +
                                 ((m & 2) ? (s01[c] >> 16) : 0) +
                                 ((m & 4) ? int32_t(int16_t(s23[c])) : 0) +
                                 ((m & 8) ? (s23[c] >> 16) : 0);
            // term index kept opaque here, so that the compiler forms the table
            // addresses in the loop instead of holding them (64-bit each) across it
            uint32_t tf = t0[c];
            asm volatile("" : "+v"(tf));
            // term-order fold: first value, then + (accel.rs:191-208)
            double2 v = kLean ? signed_coeff(ld2(mem.tc + tf), parity_and<kNarrow>(r, mem.tz[tf]))
                              : signed_coeff(c0[c], parity_and<kNarrow>(r, z0[c]));
            if (kHeavy && hs[c] >= 0) {
                v = hv[(uint32_t(hs[c]) << (tile_log2 + fold_log2)) | (sub << tile_log2) |
                       (uint32_t(m) << rs_log2) | ro];
            } else if (t1[c] > tf + 1) {
                const double2 e = kHold2 ? signed_coeff(c1[c], parity_and<kNarrow>(r, z1[c]))
                                         : signed_coeff(ld2(mem.tc + tf + 1),
                                                        parity_and<kNarrow>(r, mem.tz[tf + 1]));
                v.x = v.x + e.x;
                v.y = v.y + e.y;
                uint32_t u0 = 0;
                if (kHold4) {
                    // terms 2 and 3 from registers
                    if (t1[c] > tf + 2) {
                        const double2 e2 = signed_coeff(c2[c], parity_and<kNarrow>(r, z2[c]));
                        v.x = v.x + e2.x;
                        v.y = v.y + e2.y;
                    }
                    if (t1[c] > tf + 3) {
                        const double2 e3 = signed_coeff(c3[c], parity_and<kNarrow>(r, z3[c]));
                        v.x = v.x + e3.x;
                        v.y = v.y + e3.y;
                    }
                    u0 = 2;
                    if (kHold6) {
                        if (t1[c] > tf + 4) {
                            const double2 e4 = signed_coeff(c4[c], parity_and<kNarrow>(r, z4[c]));
                            v.x = v.x + e4.x;
                            v.y = v.y + e4.y;
                        }
                        if (t1[c] > tf + 5) {
                            const double2 e5 = signed_coeff(c5[c], parity_and<kNarrow>(r, z5[c]));
                            v.x = v.x + e5.x;
                            v.y = v.y + e5.y;
                        }
                        u0 = 4;
                    }
                }
                // the remaining terms, in term order: from the transposed tail
                // table while it has them -- the lanes, one group each, then
                // read one contiguous run per position (u counts iterations,
                // so the choice is the same for every active lane) -- else
                // from the group's own term range.  Four loads in flight were
                // slower than this serial loop, from either table.
                for (uint32_t t = tf + 2 + u0, u = u0; t < t1[c]; ++t, ++u) {
                    double2 ct;
                    if (u < op.ntail) {
                        const uint32_t slot = u * op.gstride + j;
                        ct = signed_coeff(ld2(mem.ttc + slot), parity_and<kNarrow>(r, mem.ttz[slot]));
                    } else {
                        ct = signed_coeff(ld2(mem.tc + t), parity_and<kNarrow>(r, mem.tz[t]));
                    }
                    v.x = v.x + ct.x;
                    v.y = v.y + ct.y;
                }
            }
            const uint64_t col = r ^ uint64_t(xj[c]);
            if (kMatvec) {
                const double2 xb = ld2(mem.x + col);
                acc.x = fma(v.x, xb.x, acc.x);
                acc.x = fma(-v.y, xb.y, acc.x);
                acc.y = fma(v.x, xb.y, acc.y);
                acc.y = fma(v.y, xb.x, acc.y);
            }
            if (!kMatvec || op.indices) {
                const uint32_t slot = (ro * G) + uint32_t(rank - off);
                si[slot] = col;
                sd[slot] = v;
            }
        }
        __syncthreads();
        const uint64_t b0 = tile_base + (uint64_t(m) << rs_log2);
        const uint64_t lo_r = b0 > row_begin ? b0 : row_begin;
        const uint64_t b1 = b0 + (uint64_t(1) << rs_log2);
        const uint64_t hi_r = b1 < row_end ? b1 : row_end;
        if (kMatvec) {
            // y of row ro: a fixed tree over the row's lanes
            const uint64_t rr = b0 + ro;
            const uint32_t wl = jw_log2 < 6 ? jw_log2 : 6;
            for (uint32_t off = (1u << wl) >> 1; off > 0; off >>= 1) {
                acc.x += __shfl_xor(acc.x, int(off), 64);
                acc.y += __shfl_xor(acc.y, int(off), 64);
            }
            if (jw_log2 <= 6) {
                if (lane == 0 && rr >= row_begin && rr < row_end) mem.y[rr - row_begin] = dvec2{acc.x, acc.y};
            } else if ((tid & 63) == 0) {
                red[buf * 16 + (tid >> 6)] = acc;
            }
        }
        if (op.indices && lo_r < hi_r) {
            // consecutive rows are consecutive in the output unless this is a
            // sub-operator, which always has one row per batch
            const uint32_t count = uint32_t(hi_r - lo_r) * G;
            const uint32_t src = uint32_t(lo_r - b0) * G;
            const uint64_t dst = kSub ? (lo_r - row_begin) * op.gstride + uint32_t(off)
                                      : (lo_r - row_begin) * G;
            for (uint32_t i = tid; i < count; i += kStageThreads) {
                mem.indices[dst + i] = si[src + i];
                const double2 dv = sd[src + i];
                const dvec2 vv = {dv.x, dv.y};
                mem.data[dst + i] = vv;
            }
        }
        if (kMatvec && jw_log2 > 6) {
            // rows whose JW lanes span several waves: add the wave partials
            __syncthreads();
            if (tid < (1u << rs_log2)) {
                const uint32_t wpr = 1u << (jw_log2 - 6);
                double2 sum = make_double2(0.0, 0.0);
                for (uint32_t w = 0; w < wpr; ++w) {
                    const double2 p = red[buf * 16 + tid * wpr + w];
                    sum.x += p.x;
                    sum.y += p.y;
                }
                const uint64_t rr = b0 + tid;
                if (rr >= row_begin && rr < row_end) mem.y[rr - row_begin] = dvec2{sum.x, sum.y};
            }
        } else if (!op.dbuf) {
            __syncthreads();
        }
    }
}

}  // namespace qr
