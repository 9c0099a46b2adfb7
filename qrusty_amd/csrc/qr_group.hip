// qr_group.hip -- an operator's term tables built on the device: label parse,
// make_params, grouping by x mask, rank lists.
//
// Replaces, for operators with many terms, the host steps of qr_op_create:
// Pauli::parse_label / x_indices / z_indices / phase (qrusty/src/lib.rs:
// 125-181), accel::rowwise::make_params (accel.rs:141-157), and the stable
// grouping of the terms by x mask that make_row's stable sort implies
// (accel.rs:188-208).  The results are bit-identical to the host path:
//  * k_parse_labels: one thread per label, the grammar
//    ^([+-]?)1?([ij]?)([IXYZ]+)$, qubit q = label[len-1-q], x bit for X|Y,
//    z bit for Y|Z, c' = M[p] c with the num-complex product (no fused
//    multiply-add);
//  * a stable LSD radix sort of the term indices by x, 8 bits per pass
//    (k_radix_hist, k_scan_u32, k_radix_scatter): each wavefront ranks its 64
//    keys among equal digits with eight ballots (the lanes whose digit
//    matches in every bit), waves and rounds are ordered through LDS
//    counters, so equal masks keep their input order;
//  * a segmented scan over the sorted masks: head flags, their exclusive scan
//    (group index of every term), group masks and term ranges, the grouped
//    z masks and coefficients, and each group's coefficient kind;
//  * the rank lists: one thread per group walks the mask trie from the root
//    (the same splits as the host's build_rank_lists), first counting, then
//    writing its (bit, sigma) entries.
#include "qr_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace qr {

constexpr int kGrpThreads = 256;
constexpr uint32_t kSortRounds = 16;                       // rounds of 256 keys per block
constexpr uint32_t kSortTile = kSortRounds * kGrpThreads;  // keys per block and pass
constexpr uint64_t kDeviceGroupMaxTerms = uint64_t(1) << 28;

// status of a label: 0 ok, 1 malformed, 2 more than 63 qubits
__global__ __launch_bounds__(kGrpThreads) void k_parse_labels(
    uint64_t T, const char *__restrict__ bytes, const uint64_t *__restrict__ offs,
    const double *__restrict__ coeffs, int32_t conv, uint64_t *__restrict__ x,
    uint64_t *__restrict__ z, double2 *__restrict__ cp, uint32_t *__restrict__ nq,
    uint8_t *__restrict__ status) {
#pragma clang fp contract(off)
    const uint64_t t = uint64_t(blockIdx.x) * kGrpThreads + threadIdx.x;
    if (t >= T) return;
    const char *s = bytes + offs[t];
    const uint64_t len = offs[t + 1] - offs[t];
    uint64_t i = 0;
    bool neg = false, imag = false;
    if (i < len && (s[i] == '+' || s[i] == '-')) { neg = s[i] == '-'; ++i; }
    if (i < len && s[i] == '1') ++i;
    if (i < len && (s[i] == 'i' || s[i] == 'j')) { imag = true; ++i; }
    const uint64_t n = len - i;
    uint8_t st = n == 0 ? 1 : 0;
    for (uint64_t k = i; k < len && !st; ++k)
        if (s[k] != 'I' && s[k] != 'X' && s[k] != 'Y' && s[k] != 'Z') st = 1;
    if (!st && n > 63) st = 2;
    status[t] = st;
    nq[t] = uint32_t(n > 0xFFFFFFFFull ? 0xFFFFFFFFull : n);
    if (st) return;
    uint64_t xm = 0, zm = 0;
    uint32_t ny = 0;
    for (uint64_t q = 0; q < n; ++q) {
        const char c = s[len - 1 - q];
        if (c == 'X' || c == 'Y') xm |= uint64_t(1) << q;
        if (c == 'Y' || c == 'Z') zm |= uint64_t(1) << q;
        if (c == 'Y') ++ny;
    }
    x[t] = xm;
    z[t] = zm;
    const uint32_t bp = (imag ? 1u : 0u) + (neg ? 2u : 0u);
    const uint32_t p = conv == QR_PHASE_ROWWISE ? (bp + ny) % 4u : (ny + 4u - bp) % 4u;
    // M[p] of accel.rs:147-152 = {1, -i, -1, +i}, times c as num-complex multiplies
    const double mr = p == 0 ? 1.0 : (p == 2 ? -1.0 : 0.0);
    const double mi = p == 1 ? -1.0 : (p == 3 ? 1.0 : 0.0);
    const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1];
    const double pr = mr * cr, qi = mi * ci, pi = mr * ci, qr = mi * cr;
    cp[t] = make_double2(pr - qi, pi + qr);
}

// First failing term: a label error, or a qubit count other than term 0's.
__global__ __launch_bounds__(kGrpThreads) void k_first_bad(uint64_t T, const uint8_t *__restrict__ status,
                                                           const uint32_t *__restrict__ nq,
                                                           unsigned long long *__restrict__ first) {
    const uint64_t t = uint64_t(blockIdx.x) * kGrpThreads + threadIdx.x;
    if (t >= T) return;
    if (status[t] || nq[t] != nq[0]) atomicMin(first, (unsigned long long)t);
}

__global__ __launch_bounds__(kGrpThreads) void k_iota(uint64_t T, uint32_t *__restrict__ v) {
    const uint64_t t = uint64_t(blockIdx.x) * kGrpThreads + threadIdx.x;
    if (t < T) v[t] = uint32_t(t);
}

// Digit counts of one block's keys, stored digit-major (hist[d * nblocks + b])
// so that one exclusive scan gives every (digit, block) its output offset.
__global__ __launch_bounds__(kGrpThreads) void k_radix_hist(uint64_t T, const uint64_t *__restrict__ keys,
                                                            uint32_t shift, uint32_t nblocks,
                                                            uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[256];
    const uint32_t tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const uint64_t base = uint64_t(blockIdx.x) * kSortTile;
    for (uint32_t i = tid; i < kSortTile; i += kGrpThreads) {
        const uint64_t p = base + i;
        if (p < T) atomicAdd(&cnt[uint32_t(keys[p] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[uint64_t(tid) * nblocks + blockIdx.x] = cnt[tid];
}

// Exclusive scan of n uint32 in one block of 1024 threads (4 values each per
// step, wave scans, a scan of the 16 wave sums, a running carry); *total
// receives the sum.
__global__ __launch_bounds__(1024) void k_scan_u32(uint64_t n, const uint32_t *__restrict__ in,
                                                   uint32_t *__restrict__ out,
                                                   uint32_t *__restrict__ total) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (uint64_t base = 0; base < n; base += 4096) {
        uint32_t v[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t idx = base + uint64_t(tid) * 4 + k;
            v[k] = idx < n ? in[idx] : 0u;
            s += v[k];
        }
        uint32_t incl = s;
#pragma unroll
        for (int off = 1; off < 64; off *= 2) {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= uint32_t(off)) incl += t;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        if (wave == 0) {
            uint32_t w = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
            for (int off = 1; off < 16; off *= 2) {
                const uint32_t t = __shfl_up(w, off, 64);
                if (lane >= uint32_t(off)) w += t;
            }
            if (lane < 16) wsum[lane] = w;
        }
        __syncthreads();
        uint32_t ex = carry_s + (wave ? wsum[wave - 1] : 0u) + incl - s;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t idx = base + uint64_t(tid) * 4 + k;
            if (idx < n) out[idx] = ex;
            ex += v[k];
        }
        __syncthreads();
        if (tid == 0) carry_s += wsum[15];
        __syncthreads();
    }
    if (tid == 0 && total) *total = carry_s;
}

// Stable scatter of one radix pass.  Keys are taken in position order: round
// r of the block covers positions base + 256 r + 64 w + lane (wave w).  A key's
// rank = the block's offset for its digit + keys of that digit in earlier
// rounds (run) + in earlier waves of this round (wcnt) + in earlier lanes of
// its wave (the lanes whose digit matches in all eight ballots).
__global__ __launch_bounds__(kGrpThreads) void k_radix_scatter(
    uint64_t T, const uint64_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t shift, uint32_t nblocks, const uint32_t *__restrict__ offs,
    uint64_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kGrpThreads / 64][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    run[tid] = offs[uint64_t(tid) * nblocks + blockIdx.x];
    const uint64_t base = uint64_t(blockIdx.x) * kSortTile;
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    for (uint32_t round = 0; round < kSortRounds; ++round) {
        // each thread clears (and later sums) the counters of digit tid
#pragma unroll
        for (uint32_t w = 0; w < kGrpThreads / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const uint64_t p = base + uint64_t(round) * kGrpThreads + tid;
        const bool valid = p < T;
        const uint64_t key = valid ? keys_in[p] : 0;
        const uint32_t val = valid ? vals_in[p] : 0u;
        const uint32_t d = uint32_t(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t r_in_wave = uint32_t(__popcll(peers & below));
        if (valid && r_in_wave == 0) wcnt[wave][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + r_in_wave;
            for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][d];
            if (pos < T) {  // always, when the offsets are consistent
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (uint32_t w = 0; w < kGrpThreads / 64; ++w) add += wcnt[w][tid];
        run[tid] += add;
    }
}

// Head flags of the sorted masks, and the terms' z masks and coefficients in
// sorted (grouped) order.
__global__ __launch_bounds__(kGrpThreads) void k_group_heads(uint64_t T, const uint64_t *__restrict__ keys,
                                                             const uint32_t *__restrict__ perm,
                                                             const uint64_t *__restrict__ z,
                                                             const double2 *__restrict__ cp,
                                                             uint32_t *__restrict__ head,
                                                             uint64_t *__restrict__ tz,
                                                             double2 *__restrict__ tc) {
    const uint64_t i = uint64_t(blockIdx.x) * kGrpThreads + threadIdx.x;
    if (i >= T) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
    const uint32_t t = perm[i];
    tz[i] = z[t];
    tc[i] = cp[t];
}

// Group masks and term ranges; each term clears the kind bits its
// coefficient rules out (bit 0: every coefficient real, bit 1: every one
// imaginary; flags start at 3).
__global__ __launch_bounds__(kGrpThreads) void k_group_tables(
    uint64_t T, uint32_t G, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ head,
    const uint32_t *__restrict__ hscan, const double2 *__restrict__ tc, uint64_t *__restrict__ gx,
    uint32_t *__restrict__ gterm, uint32_t *__restrict__ kind) {
    const uint64_t i = uint64_t(blockIdx.x) * kGrpThreads + threadIdx.x;
    if (i >= T) return;
    const uint32_t g = hscan[i] + head[i] - 1u;
    if (g >= G) return;  // never, when the scan is consistent
    if (head[i]) {
        gx[g] = keys[i];
        gterm[g] = uint32_t(i);
    }
    if (i == 0) gterm[G] = uint32_t(T);
    const double2 c = tc[i];
    const uint32_t clear = (c.y != 0.0 ? 1u : 0u) | (c.x != 0.0 ? 2u : 0u);
    if (clear) atomicAnd(&kind[g], ~clear);
}

__global__ __launch_bounds__(kGrpThreads) void k_group_kind(uint32_t G, const uint32_t *__restrict__ kind,
                                                            uint8_t *__restrict__ gkind) {
    const uint32_t g = blockIdx.x * kGrpThreads + threadIdx.x;
    if (g >= G) return;
    const uint32_t f = kind[g];
    gkind[g] = (f & 1u) ? 1 : ((f & 2u) ? 2 : 0);
}

__global__ __launch_bounds__(kGrpThreads) void k_fill_u32(uint64_t n, uint32_t v, uint32_t *__restrict__ a) {
    const uint64_t i = uint64_t(blockIdx.x) * kGrpThreads + threadIdx.x;
    if (i < n) a[i] = v;
}

// Rank list of group j: the splits of the mask trie on the path from the root
// to j (the highest bit in which the range's first and last masks differ,
// split where that bit turns on), +size of the right part for a left child,
// -size of the left part for a right child.  kFill = false counts them.
template <bool kFill>
__global__ __launch_bounds__(kGrpThreads) void k_rank_lists(uint32_t G, uint32_t E,
                                                            const uint64_t *__restrict__ gx,
                                                            const uint32_t *__restrict__ rl_ptr,
                                                            uint32_t *__restrict__ cnt,
                                                            RankEntry *__restrict__ rl) {
    const uint32_t j = blockIdx.x * kGrpThreads + threadIdx.x;
    if (j >= G) return;
    uint32_t lo = 0, hi = G, n = 0;
    uint32_t e = kFill ? rl_ptr[j] : 0u;
    while (hi - lo > 1) {
        const int b = 63 - __clzll(gx[lo] ^ gx[hi - 1]);
        const uint64_t bit = uint64_t(1) << b;
        uint32_t a = lo, c = hi;  // first index of [lo, hi) with the bit set
        while (a < c) {
            const uint32_t m = a + (c - a) / 2;
            if (gx[m] & bit) c = m;
            else a = m + 1;
        }
        if (kFill && e < E) rl[e++] = RankEntry{b, j < a ? int32_t(hi - a) : -int32_t(a - lo)};
        if (j < a) hi = a;
        else lo = a;
        ++n;
    }
    if (!kFill) cnt[j] = n;
}

namespace {

// device buffers of one build, freed on every exit path
struct DevBufs {
    std::vector<void *> ptrs;
    template <typename T>
    int alloc(T **p, size_t count) {
        *p = nullptr;
        QR_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T)));
        ptrs.push_back(*p);
        return QR_OK;
    }
    void release(void *p) {  // hand a buffer over (no longer freed here)
        for (auto &q : ptrs)
            if (q == p) q = nullptr;
    }
    ~DevBufs() {
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
};

inline unsigned grid_of(uint64_t n) { return unsigned((n + kGrpThreads - 1) / kGrpThreads); }

}  // namespace

int group_terms_device(const char *const *labels, const double *coeffs, uint64_t T, int32_t conv,
                       hipStream_t s, qr_op **out) {
    if (T > kDeviceGroupMaxTerms)
        return fail(QR_ERR_SIZE, "device grouping takes at most 2^28 terms");
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QR_ERR_ARG, "device index out of range");
    // labels packed as bytes with offsets (no parsing on the host); a null
    // label is packed empty, so that it fails where the host path would
    // report it (the first failing term wins)
    std::vector<uint64_t> offs(T + 1);
    offs[0] = 0;
    uint64_t first_null = T;
    for (uint64_t t = 0; t < T; ++t) {
        if (!labels[t] && first_null == T) first_null = t;
        offs[t + 1] = offs[t] + (labels[t] ? std::strlen(labels[t]) : 0);
    }
    std::vector<char> bytes(std::max<uint64_t>(offs[T], 1));
    for (uint64_t t = 0; t < T; ++t)
        if (labels[t]) std::memcpy(bytes.data() + offs[t], labels[t], offs[t + 1] - offs[t]);

    DevBufs B;
    char *d_bytes;
    uint64_t *d_offs, *d_x, *d_z, *d_k0, *d_k1, *d_tz, *d_gx = nullptr;
    double *d_coef;
    double2 *d_cp, *d_tc;
    uint32_t *d_nq, *d_v0, *d_v1, *d_head, *d_hscan, *d_tot, *d_kind = nullptr, *d_gterm = nullptr;
    uint32_t *d_cnt = nullptr, *d_rlptr = nullptr;
    uint8_t *d_status, *d_gkind = nullptr;
    unsigned long long *d_first;
    RankEntry *d_rl = nullptr;
    int rc;
    if ((rc = B.alloc(&d_bytes, bytes.size())) || (rc = B.alloc(&d_offs, T + 1)) ||
        (rc = B.alloc(&d_coef, 2 * T)) || (rc = B.alloc(&d_x, T)) || (rc = B.alloc(&d_z, T)) ||
        (rc = B.alloc(&d_cp, T)) || (rc = B.alloc(&d_nq, T)) || (rc = B.alloc(&d_status, T)) ||
        (rc = B.alloc(&d_first, 1)) || (rc = B.alloc(&d_k0, T)) || (rc = B.alloc(&d_k1, T)) ||
        (rc = B.alloc(&d_v0, T)) || (rc = B.alloc(&d_v1, T)) || (rc = B.alloc(&d_head, T)) ||
        (rc = B.alloc(&d_hscan, T)) || (rc = B.alloc(&d_tot, 2)) || (rc = B.alloc(&d_tz, T)) ||
        (rc = B.alloc(&d_tc, T)))
        return rc;
    QR_HIP_CHECK(hipMemcpyAsync(d_bytes, bytes.data(), bytes.size(), hipMemcpyHostToDevice, s));
    QR_HIP_CHECK(hipMemcpyAsync(d_offs, offs.data(), (T + 1) * 8, hipMemcpyHostToDevice, s));
    QR_HIP_CHECK(hipMemcpyAsync(d_coef, coeffs, 2 * T * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_parse_labels, dim3(grid_of(T)), dim3(kGrpThreads), 0, s, T, d_bytes, d_offs,
                       d_coef, conv, d_x, d_z, d_cp, d_nq, d_status);
    const unsigned long long none = ~0ull;
    QR_HIP_CHECK(hipMemcpy(d_first, &none, 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_first_bad, dim3(grid_of(T)), dim3(kGrpThreads), 0, s, T, d_status, d_nq, d_first);
    QR_HIP_CHECK(hipGetLastError());
    unsigned long long first = 0;
    uint32_t nq0 = 0;
    QR_HIP_CHECK(hipMemcpyAsync(&first, d_first, 8, hipMemcpyDeviceToHost, s));
    QR_HIP_CHECK(hipMemcpyAsync(&nq0, d_nq, 4, hipMemcpyDeviceToHost, s));
    QR_HIP_CHECK(hipStreamSynchronize(s));
    if (first != none) {
        // the error the host path reports first: the failing label's own
        // error, else the qubit-count mismatch
        if (first == first_null) return fail(QR_ERR_ARG, "null label");
        uint8_t st = 0;
        QR_HIP_CHECK(hipMemcpy(&st, d_status + first, 1, hipMemcpyDeviceToHost));
        if (st == 1) return fail(QR_ERR_LABEL, "error: malformed label");
        if (st == 2) return fail(QR_ERR_SIZE, "Pauli labels of more than 63 qubits are not supported");
        return fail(QR_ERR_TERMS, "SparsePauliOp::new: all supplied paulis must have the same #qubits");
    }

    // stable LSD radix sort of the term indices by x, 8 bits per pass
    const uint32_t nblocks = uint32_t((T + kSortTile - 1) / kSortTile);
    uint32_t *d_hist, *d_hoff;
    if ((rc = B.alloc(&d_hist, uint64_t(256) * nblocks)) || (rc = B.alloc(&d_hoff, uint64_t(256) * nblocks)))
        return rc;
    QR_HIP_CHECK(hipMemcpyAsync(d_k0, d_x, T * 8, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_iota, dim3(grid_of(T)), dim3(kGrpThreads), 0, s, T, d_v0);
    uint64_t *kin = d_k0, *kout = d_k1;
    uint32_t *vin = d_v0, *vout = d_v1;
    for (uint32_t shift = 0; shift < nq0; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kGrpThreads), 0, s, T, kin, shift, nblocks, d_hist);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, uint64_t(256) * nblocks, d_hist, d_hoff,
                           d_tot);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kGrpThreads), 0, s, T, kin, vin, shift,
                           nblocks, d_hoff, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // segmented scan over the sorted masks
    hipLaunchKernelGGL(k_group_heads, dim3(grid_of(T)), dim3(kGrpThreads), 0, s, T, kin, vin, d_z, d_cp,
                       d_head, d_tz, d_tc);
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, T, d_head, d_hscan, d_tot);
    QR_HIP_CHECK(hipGetLastError());
    uint32_t G = 0;
    QR_HIP_CHECK(hipMemcpyAsync(&G, d_tot, 4, hipMemcpyDeviceToHost, s));
    QR_HIP_CHECK(hipStreamSynchronize(s));
    if ((rc = B.alloc(&d_gx, G)) || (rc = B.alloc(&d_gterm, uint64_t(G) + 1)) || (rc = B.alloc(&d_kind, G)) ||
        (rc = B.alloc(&d_gkind, G)) || (rc = B.alloc(&d_cnt, G)) || (rc = B.alloc(&d_rlptr, uint64_t(G) + 1)))
        return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(G)), dim3(kGrpThreads), 0, s, uint64_t(G), 3u, d_kind);
    hipLaunchKernelGGL(k_group_tables, dim3(grid_of(T)), dim3(kGrpThreads), 0, s, T, G, kin, d_head, d_hscan,
                       d_tc, d_gx, d_gterm, d_kind);
    hipLaunchKernelGGL(k_group_kind, dim3(grid_of(G)), dim3(kGrpThreads), 0, s, G, d_kind, d_gkind);
    // rank lists: count, scan, fill
    hipLaunchKernelGGL(k_rank_lists<false>, dim3(grid_of(G)), dim3(kGrpThreads), 0, s, G, 0u, d_gx,
                       nullptr, d_cnt, nullptr);
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, uint64_t(G), d_cnt, d_rlptr, d_tot + 1);
    QR_HIP_CHECK(hipGetLastError());
    uint32_t E = 0;
    QR_HIP_CHECK(hipMemcpyAsync(&E, d_tot + 1, 4, hipMemcpyDeviceToHost, s));
    QR_HIP_CHECK(hipStreamSynchronize(s));
    QR_HIP_CHECK(hipMemcpy(d_rlptr + G, &E, 4, hipMemcpyHostToDevice));
    if ((rc = B.alloc(&d_rl, E))) return rc;
    hipLaunchKernelGGL(k_rank_lists<true>, dim3(grid_of(G)), dim3(kGrpThreads), 0, s, G, E, d_gx,
                       d_rlptr, nullptr, d_rl);
    QR_HIP_CHECK(hipGetLastError());

    // host copies of every table (the host side plans launches from them)
    qr_op *op = new (std::nothrow) qr_op;
    if (!op) return fail(QR_ERR_NOMEM, "out of host memory");
    op->nq = nq0;
    op->T = T;
    op->G = G;
    op->x.resize(T);
    op->z.resize(T);
    op->cp.resize(2 * T);
    op->gx.resize(G);
    op->gterm.resize(uint64_t(G) + 1);
    op->tz.resize(T);
    op->tc.resize(2 * T);
    op->gkind.resize(G);
    op->rl_ptr.resize(uint64_t(G) + 1);
    op->rl.resize(E);
    hipError_t e = hipSuccess;
    auto get = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    get(op->x.data(), d_x, T * 8);
    get(op->z.data(), d_z, T * 8);
    get(op->cp.data(), d_cp, T * 16);
    get(op->gx.data(), d_gx, uint64_t(G) * 8);
    get(op->gterm.data(), d_gterm, (uint64_t(G) + 1) * 4);
    get(op->tz.data(), d_tz, T * 8);
    get(op->tc.data(), d_tc, T * 16);
    get(op->gkind.data(), d_gkind, G);
    get(op->rl_ptr.data(), d_rlptr, (uint64_t(G) + 1) * 4);
    get(op->rl.data(), d_rl, uint64_t(E) * sizeof(RankEntry));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        delete op;
        return hip_fail(e, "copying the device-built tables");
    }
    finish_op(op);
    // the device tables of this device are the ones just built
    DeviceTables &t = op->dev[dev];
    t.gx = d_gx;
    t.gterm = d_gterm;
    t.tz = d_tz;
    t.tc = d_tc;
    t.rl_ptr = d_rlptr;
    t.rl = d_rl;
    t.gkind = d_gkind;
    for (void *p : {(void *)d_gx, (void *)d_gterm, (void *)d_tz, (void *)d_tc, (void *)d_rlptr,
                    (void *)d_rl, (void *)d_gkind})
        B.release(p);
    auto put = [&](auto **dst, const auto &v) -> int {
        using T_ = typename std::decay<decltype(v[0])>::type;
        QR_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(v.size(), 1) * sizeof(T_)));
        QR_HIP_CHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T_), hipMemcpyHostToDevice));
        return QR_OK;
    };
    if ((!op->heavy.empty() && (rc = put(&t.heavy, op->heavy))) ||
        (!op->leaf_heavy.empty() && (rc = put(&t.leaf_heavy, op->leaf_heavy))) ||
        (rc = upload_tail(op, t))) {
        free_tables(t);
        delete op;
        return rc;
    }
    t.nheavy = uint32_t(op->heavy.size());
    t.ready = true;
    *out = op;
    return QR_OK;
}

}  // namespace qr
