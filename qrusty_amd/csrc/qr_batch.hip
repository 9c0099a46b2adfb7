// qr_batch.hip -- many operators of uneven size: CSR builds plus one matvec
// each, as a few persistent kernels captured in one hipGraph, with an
// autotuner.
//
// The reference builds and applies operators one at a time
// (qrusty/src/lib.rs:573-585 then accel.rs:338-370).  Here the work of a whole
// batch is lists of (operator, row tile) items, one per register class.  A
// persistent grid of 1024-thread blocks takes items from its class's atomic
// counter; an item builds its
// rows with the shared LDS-staged tile builder (qr_tile.h) -- CSR entries
// bit-identical to the single-operator build -- and, fused, the same rows of
// y = A x from the entries still in registers.  No item depends on another,
// so every block leaves the loop as soon as the counter passes the end.
// Two forms (a tuning choice).  Fused: one kernel per register class builds
// and applies; operators whose G exceeds the largest LDS stage
// (kStageMaxGLean) are built and applied by the single-operator kernels,
// recorded in the same graph.  Split (needs every CSR): two build kernels,
// whose items are the tiles of every operator as sub-operators of at most
// kSubMaxG groups (an operator of at most kSubMaxG groups is one), then one
// matvec pass (k_batch_mv) over the written CSRs -- three compute kernels per
// step whatever the batch holds.
//
// Tuning knobs (qr_tuning): tiles per item (work granularity), resident
// blocks per CU (grid size), and the item order (input order or largest
// operator first).  qr_batch_tune times every candidate on the caller's
// buffers and keeps the fastest.
#include "qr_internal.h"
#include "qr_tile.h"

#include <algorithm>
#include <numeric>

namespace qr {

struct BatchOpDesc {
    TileOp op;
    uint64_t item_begin;  // first work item of this operator
    uint64_t tiles;       // row tiles of this operator
    uint32_t narrow;      // masks fit in 32 bits
    uint32_t nch;         // group chunks per thread
};

// One persistent kernel per register class (mask width x chunks per thread):
// items of that class only, taken from the class's own counter.
// Minimum waves per SIMD requested from the compiler: 8 means two resident
// 1024-thread blocks per CU (at most 64 VGPRs), as the single-operator build
// runs; 1 leaves the register budget alone.  Only the one-chunk build fits
// 64 VGPRs here: the two-chunk build spilled 25 registers there (its
// item loop costs a few more than the single-operator kernel), and without
// the bound the 256-operator batch measured 92.0-93.1 against 93.0-94.1 ms.
#ifndef QR_BATCH_WAVES
#define QR_BATCH_WAVES(matvec, chunks, mode) \
    ((!(matvec) && ((chunks) <= 1 || (mode) == kTileSub)) ? 8 : 1)
#endif
template <bool kNarrow, int kChunks, bool kMatvec, int kMode = kTilePlain>
__global__ __launch_bounds__(kStageThreads, QR_BATCH_WAVES(kMatvec, kChunks, kMode)) void k_batch(const BatchOpDesc *__restrict__ ops,
                                                         uint32_t nops, uint64_t total_items,
                                                         uint32_t tiles_per_item,
                                                         unsigned long long *__restrict__ counter) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ unsigned long long item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1ull);
        __syncthreads();
        const uint64_t item = item_s;
        __syncthreads();
        if (item >= total_items) break;
        // operator owning the item: last k with item_begin <= item
        uint32_t lo = 0, hi = nops;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (ops[mid].item_begin <= item) lo = mid;
            else hi = mid;
        }
        const BatchOpDesc &d = ops[lo];
        const uint64_t t0 = (item - d.item_begin) * tiles_per_item;
        const uint64_t t1 = t0 + tiles_per_item < d.tiles ? t0 + tiles_per_item : d.tiles;
        for (uint64_t t = t0; t < t1; ++t) {
            build_tile_staged<kNarrow, kChunks, kMatvec, kMode>(d.op, t, smem_raw);
            __syncthreads();  // the stage is reused by the next tile
        }
    }
}

// Split form, second pass: y_k = A_k x_k from the CSR just written, over
// (operator, block of kMvRows rows) items taken from one counter.  Operators
// of at least kMvStagedMinRow groups run the row-staged walker of qr_mv.h
// (contiguous row streams, gathers of 8 neighbouring rows per instruction);
// shorter rows 8 lanes per row, a 256-thread block covering 32 consecutive
// rows per step, so that the gathers x[r ^ x_g] of neighbouring rows share
// cache lines.  Every row of a batch operator holds exactly G entries.  The
// arrays come from descriptors in memory, so they are read through pointers
// cast to the global address space (flat instructions otherwise).
struct BatchMvDesc {
    const uint64_t *indices;
    const double2 *data;
    const double2 *x;
    double2 *y;
    uint64_t rows, G, item_begin;
};
// 128 rows per item: the 256-operator uneven batch measured 100 ms per step
// at 1024 rows, 94-95 at 512, 93.6 at 256, 92.5-94 at 128, 93 at 64 and 32
// (smaller items leave less of a tail).
constexpr uint32_t kMvRows = 128;

// Rows [r0, r1) of one operator's y = A x, 8 lanes per row (rows of fewer
// than kMvStagedMinRow entries; plain loads, which beat nontemporal ones for
// short rows as in k_csr_matvec).
__device__ __forceinline__ void batch_mv_rows(const BatchMvDesc &d, uint64_t r0, uint64_t r1,
                                              uint32_t lane) {
    auto ld = [](const auto p) { return *p; };
    const auto indices = as_global(d.indices);
    const auto data = as_global(reinterpret_cast<const dvec2 *>(d.data));
    const auto x = as_global(reinterpret_cast<const dvec2 *>(d.x));
    const auto y = as_global(reinterpret_cast<dvec2 *>(d.y));
    for (uint64_t rb = r0; rb < r1; rb += 32) {
        const uint64_t row = rb + (threadIdx.x >> 3);
        double2 acc = make_double2(0.0, 0.0);
        if (row < r1) {
            const uint64_t k1 = (row + 1) * d.G;
            uint64_t k = row * d.G + lane;
            // full chunks of 64 entries for all 8 lanes of the row together
            for (; k - lane + 64 <= k1; k += 64) {
                uint64_t c[8];
                double2 a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    c[u] = ld(indices + k + 8 * u);
                    const dvec2 av = ld(data + k + 8 * u);
                    a[u] = make_double2(av.x, av.y);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const dvec2 xb = x[c[u]];
                    acc.x = fma(a[u].x, xb.x, acc.x);
                    acc.x = fma(-a[u].y, xb.y, acc.x);
                    acc.y = fma(a[u].x, xb.y, acc.y);
                    acc.y = fma(a[u].y, xb.x, acc.y);
                }
            }
            if (k < k1) {
                // the fewer than 8 entries left (all of a short row):
                // predicated, every load issued before the first is used
                uint64_t c[8];
                double2 a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    c[u] = 0;
                    a[u] = make_double2(0.0, 0.0);
                    if (k + 8 * u < k1) {
                        c[u] = ld(indices + k + 8 * u);
                        const dvec2 av = ld(data + k + 8 * u);
                        a[u] = make_double2(av.x, av.y);
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (k + 8 * u < k1) {
                        const dvec2 xb = x[c[u]];
                        acc.x = fma(a[u].x, xb.x, acc.x);
                        acc.x = fma(-a[u].y, xb.y, acc.x);
                        acc.y = fma(a[u].x, xb.y, acc.y);
                        acc.y = fma(a[u].y, xb.x, acc.y);
                    }
                }
            }
        }
#pragma unroll
        for (int off = 4; off > 0; off /= 2) {
            acc.x += __shfl_xor(acc.x, off, 8);
            acc.y += __shfl_xor(acc.y, off, 8);
        }
        if (row < r1 && lane == 0) y[row] = dvec2{acc.x, acc.y};
    }
}

// dynamic LDS: the row-staged walker's regions for the block's 4 waves
constexpr size_t kBatchMvLds = 4 * MvStagedShape<kMvStagedChunk>::kWaveBytes;

// (3 waves per SIMD: the LDS of three blocks fills a CU)
__global__ __launch_bounds__(256, 3) void k_batch_mv(const BatchMvDesc *__restrict__ ops, uint32_t nops,
                                                  uint64_t total_items,
                                                  unsigned long long *__restrict__ counter) {
    using S = MvStagedShape<kMvStagedChunk>;
    extern __shared__ __align__(16) unsigned char mv_lds[];
    __shared__ unsigned long long item_s;
    const uint32_t lane = threadIdx.x & 7u;
    const uint32_t wave = threadIdx.x / 64;
    uint64_t *s_idx = reinterpret_cast<uint64_t *>(mv_lds + size_t(wave) * S::kWaveBytes);
    dvec2 *s_dat = reinterpret_cast<dvec2 *>(s_idx + S::kRows * S::kIdxStride);
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1ull);
        __syncthreads();
        const uint64_t item = item_s;
        __syncthreads();
        if (item >= total_items) break;
        uint32_t lo = 0, hi = nops;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (ops[mid].item_begin <= item) lo = mid;
            else hi = mid;
        }
        const BatchMvDesc &d = ops[lo];
        const uint64_t r0 = (item - d.item_begin) * kMvRows;
        const uint64_t r1 = r0 + kMvRows < d.rows ? r0 + kMvRows : d.rows;
        if (d.G >= kMvStagedMinRow) {
            const uint64_t G = d.G;
            auto bounds = [G](uint64_t r, uint64_t *k0, uint64_t *len) {
                *k0 = r * G;
                *len = G;
            };
            for (uint64_t rb = r0; rb < r1; rb += 4 * S::kRows)
                mv_staged_rows<kMvStagedChunk>(rb + wave * S::kRows, r1, d.rows, bounds, d.indices,
                                               d.data, d.x, d.y, s_idx, s_dat, 0, 0);
        } else {
            batch_mv_rows(d, r0, r1, lane);
        }
    }
}

// graph-safe reset of the work counter (a kernel, not a memset node)
__global__ void k_zero_counter(unsigned long long *counter) { *counter = 0ull; }

}  // namespace qr

using namespace qr;

struct qr_batch {
    std::vector<const qr_op *> ops;
    int dev = -1;
    hipStream_t cap = nullptr;  // private capture stream (default priority)
    BatchOpDesc *d_desc = nullptr;
    size_t d_desc_cap = 0;
    BatchMvDesc *d_mv = nullptr;
    size_t d_mv_cap = 0;
    unsigned long long *d_counter = nullptr;
    int ncu = 0;
    // graph of the last (buffers, tuning)
    bool have_graph = false;
    std::vector<uintptr_t> graph_key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool have_tuned = false;
    qr_tuning tuned{};
};

namespace {

// register classes: group chunks per thread in {1, 2, kStageMaxChunks,
// kStageLeanChunks} for whole operators, and {1, 2} for the sub-operators
// (classes 4 and 5) of the split form's operators of more than kSubMaxG
// groups.  A batch operator builds all 2^n rows, so n <= 32 and masks are
// 32-bit.
constexpr int kClasses = 6;
constexpr int kSubClass = 4;
inline int class_of(uint32_t nch) {
    return nch <= 1 ? 0 : nch <= 2 ? 1 : nch <= uint32_t(kStageMaxChunks) ? 2 : 3;
}

struct Plan {
    bool split = false;                       // build pass, then k_batch_mv
    std::vector<BatchMvDesc> mv;              // split form: one per staged operator
    uint64_t mv_items = 0;
    std::vector<BatchOpDesc> desc;            // grouped by class, in item order
    uint32_t cls_begin[kClasses + 1] = {};    // descriptor range of each class
    uint64_t cls_items[kClasses] = {};
    size_t cls_shmem[kClasses] = {};
    std::vector<uint32_t> big;  // operators built by the single-operator kernels
};

int make_plan(qr_batch *b, uint64_t *const *d_indptr, uint64_t *const *d_indices,
              double *const *d_data, const double *const *d_x, double *const *d_y,
              const qr_tuning &tu, Plan &plan) {
    const uint32_t nops = uint32_t(b->ops.size());
    std::vector<uint32_t> order(nops);
    std::iota(order.begin(), order.end(), 0u);
    if (tu.order == 1)  // largest operator first
        std::stable_sort(order.begin(), order.end(), [b](uint32_t a, uint32_t c) {
            return (b->ops[a]->G << b->ops[a]->nq) > (b->ops[c]->G << b->ops[c]->nq);
        });
    const uint32_t tpi = tu.tiles_per_item ? tu.tiles_per_item : 1;
    // the split form needs every staged operator's CSR
    plan.split = tu.split != 0;
    for (uint32_t k = 0; plan.split && k < nops; ++k)
        if (!(d_indices && d_indices[k] && d_data && d_data[k]))
            return fail(QR_ERR_ARG, "the split batch form needs CSR output buffers for every operator");
    std::vector<BatchOpDesc> per[kClasses];
    auto add_item = [&](int c, const TileOp &top, uint64_t tiles, size_t shmem) {
        BatchOpDesc d{};
        d.op = top;
        d.tiles = tiles;
        d.narrow = 1;
        d.nch = top.nchunks;
        d.item_begin = plan.cls_items[c];
        plan.cls_items[c] += (tiles + tpi - 1) / tpi;
        plan.cls_shmem[c] = std::max(plan.cls_shmem[c], shmem);
        per[c].push_back(d);
    };
    for (uint32_t k : order) {
        const qr_op *op = b->ops[k];
        if (op->nq > 32) return fail(QR_ERR_SIZE, "batch operators are limited to 32 qubits");
        if (!d_x || !d_x[k] || !d_y || !d_y[k]) return fail(QR_ERR_ARG, "null state pointer in batch");
        const bool has_csr = d_indices && d_indices[k] && d_data && d_data[k];
        DeviceTables *t;
        int rc = device_tables(op, &t);
        if (rc) return rc;
        const uint32_t G = uint32_t(op->G);
        const uint64_t N = uint64_t(1) << op->nq;
        uint64_t *ip = d_indptr ? d_indptr[k] : nullptr;
        auto *dat = has_csr ? reinterpret_cast<double2 *>(d_data[k]) : nullptr;
        auto *xk = reinterpret_cast<const double2 *>(d_x[k]);
        auto *yk = reinterpret_cast<double2 *>(d_y[k]);
        if (plan.split) {
            // split form: every operator in the two sub-operator classes (1 or
            // 2 group chunks per thread).  One of at most kSubMaxG groups is a
            // sub-operator of itself (gbase = nanc = 0, gstride = G); a larger
            // one is its subtrees of the mask trie, each filling a contiguous
            // segment of every row, the first of which writes indptr.  Heavy
            // groups are folded inline.  The matvec pass then reads the CSR.
            if (op->leaves.empty()) {
                const uint32_t jw = stage_jw_log2(G), nch = (G + (1u << jw) - 1) >> jw;
                uint32_t dbuf = 0;
                const size_t sb = stage_bytes(G, jw, &dbuf);
                add_item(kSubClass + (nch <= 1 ? 0 : 1),
                         TileOp{t->gx, t->gterm, t->tz, t->tc, t->rl_ptr, t->rl, d_indices[k], dat, ip, xk,
                                yk, 0, N, G, jw, nch, dbuf, nullptr, 0, 0, 0, G},
                         ((N - 1) >> (10 - jw + 4)) + 1, sb);
            } else {
                for (const auto &lf : op->leaves) {
                    const uint32_t g = lf.s1 - lf.s0, nch = (g + 1023) >> 10;
                    uint32_t dbuf = 0;
                    const size_t sb = stage_bytes(g, 10, &dbuf);
                    add_item(kSubClass + (nch <= 1 ? 0 : 1),
                             TileOp{t->gx + lf.s0, t->gterm + lf.s0, t->tz, t->tc, t->rl_ptr + lf.s0, t->rl,
                                    d_indices[k], dat, lf.s0 == 0 ? ip : nullptr, xk, yk, 0, N, g, 10, nch,
                                    dbuf, nullptr, 0, lf.s0, lf.depth, G},
                             ((N - 1) >> 4) + 1, sb);
                }
            }
            plan.mv.push_back(BatchMvDesc{d_indices[k], dat, xk, yk, N, G, plan.mv_items});
            plan.mv_items += (N + kMvRows - 1) / kMvRows;
            continue;
        }
        if (op->G > kStageMaxGLean) {
            // fused form, more than kStageMaxGLean groups: built and applied
            // by the single-operator kernels
            if (!has_csr || !ip)
                return fail(QR_ERR_ARG, "batch operators with more than 6144 distinct x masks "
                                        "need CSR output buffers");
            plan.big.push_back(k);
            continue;
        }
        // fused form: the whole operator in the class of its chunk count
        const uint32_t jw = stage_jw_log2(G), nch = (G + (1u << jw) - 1) >> jw;
        uint32_t dbuf = 0;
        const size_t sb = stage_bytes(G, jw, &dbuf);
        add_item(class_of(nch),
                 TileOp{t->gx, t->gterm, t->tz, t->tc, t->rl_ptr, t->rl, has_csr ? d_indices[k] : nullptr, dat,
                        ip, xk, yk, 0, N, G, jw, nch, dbuf, nullptr, 0, 0, 0, G},
                 ((N - 1) >> (10 - jw + 4)) + 1, sb);
    }
    for (int c = 0; c < kClasses; ++c) {
        plan.cls_begin[c] = uint32_t(plan.desc.size());
        plan.desc.insert(plan.desc.end(), per[c].begin(), per[c].end());
    }
    plan.cls_begin[kClasses] = uint32_t(plan.desc.size());
    return QR_OK;
}

template <bool kNarrow, int kChunks, int kMode = kTilePlain>
static void launch_class(qr_batch *b, const Plan &plan, int c, uint32_t grid_cap, uint32_t tpi,
                         hipStream_t s) {
    const uint64_t items = plan.cls_items[c];
    if (items == 0) return;
    const uint64_t grid = std::min<uint64_t>(grid_cap, items);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, b->d_counter + c);
    if (plan.split || kMode != kTilePlain)
        hipLaunchKernelGGL((k_batch<kNarrow, kChunks, false, kMode>), dim3(unsigned(grid)),
                           dim3(kStageThreads), plan.cls_shmem[c], s, b->d_desc + plan.cls_begin[c],
                           plan.cls_begin[c + 1] - plan.cls_begin[c], items, tpi,
                           b->d_counter + c);
    else if constexpr (kMode == kTilePlain)
        hipLaunchKernelGGL((k_batch<kNarrow, kChunks, true>), dim3(unsigned(grid)),
                           dim3(kStageThreads), plan.cls_shmem[c], s, b->d_desc + plan.cls_begin[c],
                           plan.cls_begin[c + 1] - plan.cls_begin[c], items, tpi,
                           b->d_counter + c);
}

int record(qr_batch *b, const Plan &plan, const qr_tuning &tu, uint64_t *const *d_indptr,
           uint64_t *const *d_indices, double *const *d_data, const double *const *d_x,
           double *const *d_y, hipStream_t s) {
    const uint32_t bpc = tu.blocks_per_cu ? tu.blocks_per_cu : 1;
    const uint32_t cap = uint32_t(b->ncu) * bpc;
    const uint32_t tpi = tu.tiles_per_item ? tu.tiles_per_item : 1u;
    launch_class<true, 1>(b, plan, 0, cap, tpi, s);
    launch_class<true, 2>(b, plan, 1, cap, tpi, s);
    launch_class<true, kStageMaxChunks>(b, plan, 2, cap, tpi, s);
    launch_class<true, kStageLeanChunks>(b, plan, 3, cap, tpi, s);
    if (plan.split) {
        launch_class<true, 1, kTileSub>(b, plan, kSubClass, cap, tpi, s);
        launch_class<true, 2, kTileSub>(b, plan, kSubClass + 1, cap, tpi, s);
    }
    QR_HIP_CHECK(hipGetLastError());
    // fused form: operators of more than kStageMaxGLean groups, built and
    // applied by the single-operator kernels
    for (uint32_t k : plan.big) {
        const qr_op *op = b->ops[k];
        DeviceTables *t;
        int rc = device_tables(op, &t);
        if (rc) return rc;
        const uint64_t N = uint64_t(1) << op->nq;
        rc = launch_csr_build(op, t, 0, N, d_indptr[k], d_indices[k], d_data[k], s);
        if (rc) return rc;
        rc = launch_csr_matvec(N, N, N * op->G, d_indptr[k], d_indices[k], d_data[k], d_x[k], d_y[k], s);
        if (rc) return rc;
    }
    if (plan.split && plan.mv_items) {
        // 8 resident 256-thread blocks per CU and slot
        const uint64_t grid = std::min<uint64_t>(uint64_t(b->ncu) * 8 * bpc, plan.mv_items);
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, b->d_counter + kClasses);
        hipLaunchKernelGGL(k_batch_mv, dim3(unsigned(grid)), dim3(256), kBatchMvLds, s, b->d_mv,
                           uint32_t(plan.mv.size()), plan.mv_items, b->d_counter + kClasses);
        QR_HIP_CHECK(hipGetLastError());
    }
    return QR_OK;
}

int prepare(qr_batch *b, const Plan &plan) {
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (b->dev != dev) {
        if (b->dev >= 0) return fail(QR_ERR_ARG, "a batch is bound to the device it first ran on");
        b->dev = dev;
        QR_HIP_CHECK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        QR_HIP_CHECK(hipMalloc(&b->d_counter, (kClasses + 1) * sizeof(unsigned long long)));
        hipDeviceProp_t prop;
        QR_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        b->ncu = prop.multiProcessorCount;
        const int maxsh = int(stage_bytes(kStageMaxGLean, 10, nullptr));
        // every kernel launch_class may launch (the sub-operator classes run
        // in the split form only)
        const void *kerns[] = {
            reinterpret_cast<const void *>(k_batch<true, 1, true>),
            reinterpret_cast<const void *>(k_batch<true, 2, true>),
            reinterpret_cast<const void *>(k_batch<true, kStageMaxChunks, true>),
            reinterpret_cast<const void *>(k_batch<true, kStageLeanChunks, true>),
            reinterpret_cast<const void *>(k_batch<true, 1, false>),
            reinterpret_cast<const void *>(k_batch<true, 2, false>),
            reinterpret_cast<const void *>(k_batch<true, kStageMaxChunks, false>),
            reinterpret_cast<const void *>(k_batch<true, kStageLeanChunks, false>),
            reinterpret_cast<const void *>(k_batch<true, 1, false, kTileSub>),
            reinterpret_cast<const void *>(k_batch<true, 2, false, kTileSub>)};
        for (const void *kf : kerns)
            QR_HIP_CHECK(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, maxsh));
    }
    // a previous graph or tuning pass may still read the descriptors
    QR_HIP_CHECK(hipDeviceSynchronize());
    if (b->d_desc_cap < plan.desc.size()) {
        if (b->d_desc) (void)hipFree(b->d_desc);
        b->d_desc = nullptr;
        b->d_desc_cap = 0;
        QR_HIP_CHECK(hipMalloc(&b->d_desc, std::max<size_t>(plan.desc.size(), 1) * sizeof(BatchOpDesc)));
        b->d_desc_cap = plan.desc.size();
    }
    if (!plan.desc.empty())
        QR_HIP_CHECK(hipMemcpy(b->d_desc, plan.desc.data(), plan.desc.size() * sizeof(BatchOpDesc),
                               hipMemcpyHostToDevice));
    if (b->d_mv_cap < plan.mv.size()) {
        if (b->d_mv) (void)hipFree(b->d_mv);
        b->d_mv = nullptr;
        b->d_mv_cap = 0;
        QR_HIP_CHECK(hipMalloc(&b->d_mv, plan.mv.size() * sizeof(BatchMvDesc)));
        b->d_mv_cap = plan.mv.size();
    }
    if (!plan.mv.empty())
        QR_HIP_CHECK(hipMemcpy(b->d_mv, plan.mv.data(), plan.mv.size() * sizeof(BatchMvDesc),
                               hipMemcpyHostToDevice));
    return QR_OK;
}

void drop_graph(qr_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->have_graph = false;
}

std::vector<uintptr_t> key_of(const qr_batch *b, uint64_t *const *d_indptr,
                              uint64_t *const *d_indices, double *const *d_data,
                              const double *const *d_x, double *const *d_y, const qr_tuning &tu) {
    std::vector<uintptr_t> k;
    const size_t n = b->ops.size();
    k.reserve(5 * n + 3);
    for (size_t i = 0; i < n; ++i) {
        k.push_back(d_indptr ? uintptr_t(d_indptr[i]) : 0);
        k.push_back(d_indices ? uintptr_t(d_indices[i]) : 0);
        k.push_back(d_data ? uintptr_t(d_data[i]) : 0);
        k.push_back(uintptr_t(d_x[i]));
        k.push_back(uintptr_t(d_y[i]));
    }
    k.push_back(tu.tiles_per_item);
    k.push_back(tu.blocks_per_cu);
    k.push_back(tu.order);
    k.push_back(tu.split);
    return k;
}

const qr_tuning kDefaultTuning{1, 1, 1, 0};

}  // namespace

extern "C" {

int qr_batch_create(const qr_op *const *ops, uint32_t nops, qr_batch **out) {
    if (!out) return fail(QR_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!ops || nops == 0) return fail(QR_ERR_ARG, "empty batch");
    for (uint32_t k = 0; k < nops; ++k)
        if (!ops[k]) return fail(QR_ERR_ARG, "null operator in batch");
    qr_batch *b = new (std::nothrow) qr_batch;
    if (!b) return fail(QR_ERR_NOMEM, "out of host memory");
    b->ops.assign(ops, ops + nops);
    *out = b;
    return QR_OK;
}

int qr_batch_destroy(qr_batch *b) {
    if (!b) return QR_OK;
    drop_graph(b);
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_mv) (void)hipFree(b->d_mv);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return QR_OK;
}

int qr_batch_run(qr_batch *b, uint64_t *const *d_indptr, uint64_t *const *d_indices,
                 double *const *d_data, const double *const *d_x, double *const *d_y,
                 const qr_tuning *tuning, void *stream) {
    if (!b || !d_x || !d_y) return fail(QR_ERR_ARG, "null argument");
    const qr_tuning tu = tuning ? *tuning : (b->have_tuned ? b->tuned : kDefaultTuning);
    if (tu.tiles_per_item == 0 || tu.tiles_per_item > 4096 || tu.blocks_per_cu == 0 ||
        tu.blocks_per_cu > 8 || tu.order > 1 || tu.split > 1)
        return fail(QR_ERR_ARG, "tuning out of range");
    hipStream_t s = (hipStream_t)stream;
    std::vector<uintptr_t> key = key_of(b, d_indptr, d_indices, d_data, d_x, d_y, tu);
    if (!(b->have_graph && key == b->graph_key)) {
        Plan plan;
        int rc = make_plan(b, d_indptr, d_indices, d_data, d_x, d_y, tu, plan);
        if (rc) return rc;
        rc = prepare(b, plan);
        if (rc) return rc;
        drop_graph(b);
        QR_HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        rc = record(b, plan, tu, d_indptr, d_indices, d_data, d_x, d_y, b->cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        b->graph = g;
        QR_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->have_graph = true;
        b->graph_key = key;
    }
    QR_HIP_CHECK(hipGraphLaunch(b->exec, s));
    return QR_OK;
}

int qr_batch_tune(qr_batch *b, uint64_t *const *d_indptr, uint64_t *const *d_indices,
                  double *const *d_data, const double *const *d_x, double *const *d_y,
                  qr_tuning *best, void *stream) {
    if (!b || !d_x || !d_y) return fail(QR_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t tpis[] = {1, 2, 4, 8};
    const uint32_t bpcs[] = {1, 2};
    const uint32_t orders[] = {0, 1};
    // the split form is a candidate only when every operator keeps its CSR
    bool can_split = d_indices && d_data;
    for (size_t k = 0; can_split && k < b->ops.size(); ++k)
        can_split = d_indices[k] && d_data[k];
    hipEvent_t e0, e1;
    QR_HIP_CHECK(hipEventCreate(&e0));
    QR_HIP_CHECK(hipEventCreate(&e1));
    float best_ms = -1.0f;
    qr_tuning bt = kDefaultTuning;
    int rc = QR_OK;
    for (uint32_t split = 0; split < (can_split ? 2u : 1u); ++split)
    for (uint32_t order : orders)
        for (uint32_t bpc : bpcs)
            for (uint32_t tpi : tpis) {
                const qr_tuning tu{tpi, bpc, order, split};
                Plan plan;
                if ((rc = make_plan(b, d_indptr, d_indices, d_data, d_x, d_y, tu, plan)) ||
                    (rc = prepare(b, plan)))
                    goto done;
                // one untimed pass (first touch), then one timed pass
                for (int rep = 0; rep < 2; ++rep) {
                    if (rep == 1 && hipEventRecord(e0, s) != hipSuccess) { rc = fail(QR_ERR_HIP, "event"); goto done; }
                    if ((rc = record(b, plan, tu, d_indptr, d_indices, d_data, d_x, d_y, s))) goto done;
                }
                if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = fail(QR_ERR_HIP, "event");
                    goto done;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (best_ms < 0.0f || ms < best_ms) { best_ms = ms; bt = tu; }
            }
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    drop_graph(b);  // the descriptors now hold the last candidate's plan
    b->tuned = bt;
    b->have_tuned = true;
    if (best) *best = bt;
    return QR_OK;
}

}  // extern "C"
