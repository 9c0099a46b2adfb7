// qr_api.cpp -- host side of libqrhip: labels, operators, modes, C ABI glue.
//
// Mirrors the reference's host logic (qrusty/src/lib.rs:117-585 and
// accel.rs:141-157) and prepares the device tables the HIP kernels read:
//   * terms grouped by x mask (a stable sort, so a group's terms keep their
//     input order, which is the order make_row folds them in, accel.rs:188-208);
//   * per-group "rank lists" that give each group's position in the ascending
//     column order of any row r without sorting (see build_rank_lists).
#include "qr_internal.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <utility>

namespace qr {

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(QR_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// Label grammar of qrusty/src/lib.rs:127: ^([+-]?)1?([ij]?)([IXYZ]+)$.
// Qubit q is label[len-1-q] (lib.rs:144); x bit for X|Y and z bit for Y|Z
// (lib.rs:80-91); phase = (base_phase + #Y) mod 4 (lib.rs:179-181).
static int parse_label(const char *s, uint32_t *nq, uint32_t *base_phase, uint32_t *ny,
                       uint64_t *x, uint64_t *z) {
    if (!s) return fail(QR_ERR_ARG, "null label");
    const char *p = s;
    bool neg = false, imag = false;
    if (*p == '+' || *p == '-') { neg = (*p == '-'); ++p; }
    if (*p == '1') ++p;
    if (*p == 'i' || *p == 'j') { imag = true; ++p; }
    size_t n = std::strlen(p);
    if (n == 0) return fail(QR_ERR_LABEL, "error: malformed label");
    for (size_t k = 0; k < n; ++k)
        if (p[k] != 'I' && p[k] != 'X' && p[k] != 'Y' && p[k] != 'Z')
            return fail(QR_ERR_LABEL, "error: malformed label");
    if (n > 63) return fail(QR_ERR_SIZE, "Pauli labels of more than 63 qubits are not supported");
    uint64_t xm = 0, zm = 0;
    uint32_t y = 0;
    for (size_t q = 0; q < n; ++q) {
        char c = p[n - 1 - q];
        if (c == 'X' || c == 'Y') xm |= uint64_t(1) << q;
        if (c == 'Y' || c == 'Z') zm |= uint64_t(1) << q;
        if (c == 'Y') ++y;
    }
    *nq = uint32_t(n);
    *base_phase = (imag ? 1u : 0u) + (neg ? 2u : 0u);
    *ny = y;
    *x = xm;
    *z = zm;
    return QR_OK;
}

// Rank lists.  Sort the G distinct masks ascending (the column order of row 0)
// and view them as a binary trie, splitting at the highest differing bit.
// Row r orders its columns r ^ x ascending, which is the same trie with the
// two children of every node splitting at bit b swapped when bit b of r is
// set.  A group at sorted position j therefore sits at
//     rank_r(j) = j + sum over the splitting nodes on its path with r_b = 1 of
//                 (+size of right child if j is in the left child,
//                  -size of left child  if j is in the right child).
// Each (bit, sigma) pair is one RankEntry; a path has at most min(n, G-1) of them.
static void build_rank_lists(const std::vector<uint64_t> &gx, std::vector<uint32_t> &ptr,
                             std::vector<RankEntry> &out) {
    const size_t G = gx.size();
    std::vector<std::vector<RankEntry>> lists(G);
    struct Range { size_t lo, hi; };
    std::vector<Range> stack;
    stack.push_back({0, G});
    while (!stack.empty()) {
        Range rg = stack.back();
        stack.pop_back();
        if (rg.hi - rg.lo <= 1) continue;
        uint64_t diff = gx[rg.lo] ^ gx[rg.hi - 1];
        int b = 63 - __builtin_clzll(diff);
        uint64_t bit = uint64_t(1) << b;
        // sorted and sharing all bits above b: first index with bit b set
        size_t m = size_t(std::partition_point(gx.begin() + rg.lo, gx.begin() + rg.hi,
                                               [bit](uint64_t v) { return (v & bit) == 0; }) -
                          gx.begin());
        int32_t right = int32_t(rg.hi - m), left = int32_t(m - rg.lo);
        for (size_t j = rg.lo; j < m; ++j) lists[j].push_back({b, right});
        for (size_t j = m; j < rg.hi; ++j) lists[j].push_back({b, -left});
        stack.push_back({rg.lo, m});
        stack.push_back({m, rg.hi});
    }
    ptr.assign(G + 1, 0);
    for (size_t j = 0; j < G; ++j) ptr[j + 1] = ptr[j] + uint32_t(lists[j].size());
    out.clear();
    out.reserve(ptr[G]);
    for (size_t j = 0; j < G; ++j) out.insert(out.end(), lists[j].begin(), lists[j].end());
}

void free_tables(DeviceTables &t) {
    void *ptrs[] = {t.gx, t.gterm, t.tz, t.tc, t.rl_ptr, t.rl, t.gkind, t.heavy, t.leaf_heavy, t.ttz, t.ttc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &pt : t.pat) {
        void *pp[] = {pt.grp, pt.chunk, pt.run, pt.ug, pt.pzl, pt.ptptr, pt.ptz, pt.ptc};
        for (void *p : pp)
            if (p) (void)hipFree(p);
    }
    t = DeviceTables{};
}

// Patterns: inside a tile of 2^L rows (fixed high bits h, low bits l), term t
// of group g contributes (-1)^popcount(h & z_t) * c'_t * (-1)^popcount(l & zl_t)
// with zl_t = z_t & (2^L - 1).  Terms of a group that share zl_t therefore
// merge, per tile, into one coefficient.  The kernel splits l into tb
// thread-index bits and the row-index bits above them; a pattern whose zl
// lies in the thread bits costs once per thread, one whose zl lies in the
// row bits once per block (a small per-tile table), and only the straddling
// ones cost per row.  Terms keep input order inside a pattern.  With `pairs`
// the low bits are taken in the paired-row order (pair_row_bits).
void build_patterns(const qr_op *op, int L, int tb, uint32_t pcap, uint32_t ucap, bool staged,
                    bool pairs, HostPatterns &out) {
    const uint64_t mask = L >= 64 ? ~uint64_t(0) : ((uint64_t(1) << L) - 1);
    const uint64_t tmask = (uint64_t(1) << tb) - 1;
    out = HostPatterns{};
    out.ptptr.push_back(0);
    ApplyChunk cur{};
    const uint64_t far_mask = ~mask;
    // closes the chunk: its groups are reordered by class, or (staged) by
    // partner tile and then class (patterns stay where they are; the table
    // rows follow their groups)
    auto close_chunk = [&out, &cur, far_mask, staged, pairs]() {
        const uint32_t g0 = cur.cg[0], g1 = uint32_t(out.grp.size());
        cur.p1 = uint32_t(out.pzl.size());
        cur.u1 = uint32_t(out.ug.size());
        cur.r0 = cur.r1 = uint32_t(out.run.size());
        if (g1 > g0) {
            auto cls = [far_mask, pairs](const ApplyGroup &a) {
                return ((a.m & far_mask) ? 1u : 0u) | ((a.u >= 0 || a.ps != a.pk) ? 2u : 0u) |
                       ((a.flags & kGroupReal) ? 0u : 4u) | ((pairs && (a.m & 1)) ? 8u : 0u);
            };
            if (!staged) {
                std::stable_sort(out.grp.begin() + g0, out.grp.end(),
                                 [&cls](const ApplyGroup &a, const ApplyGroup &b) { return cls(a) < cls(b); });
                uint32_t c = 0;
                for (uint32_t g = g0; g <= g1; ++g) {
                    const uint32_t cg = g < g1 ? cls(out.grp[g]) : uint32_t(kApplyClasses);
                    while (c <= cg && c <= uint32_t(kApplyClasses)) cur.cg[c++] = g;
                }
            } else {
                // run class: bit 0 row-varying, bit 1 complex (cls without the
                // far bit; the staged form has no paired rows)
                auto key = [&cls, far_mask](const ApplyGroup &a) {
                    return std::make_pair(a.m & far_mask, (cls(a) >> 1) & 3u);
                };
                std::stable_sort(out.grp.begin() + g0, out.grp.end(),
                                 [&key](const ApplyGroup &a, const ApplyGroup &b) { return key(a) < key(b); });
                for (uint32_t g = g0; g < g1;) {
                    ApplyRun r{};
                    r.hi = out.grp[g].m & far_mask;
                    uint32_t e = g;
                    while (e < g1 && (out.grp[e].m & far_mask) == r.hi) ++e;
                    uint32_t c = 0;
                    for (uint32_t k = g; k <= e; ++k) {
                        const uint32_t ck = k < e ? key(out.grp[k]).second : uint32_t(kRunClasses);
                        while (c <= ck && c <= uint32_t(kRunClasses)) r.g[c++] = k;
                    }
                    out.run.push_back(r);
                    g = e;
                }
                cur.r1 = uint32_t(out.run.size());
                for (uint32_t c = 0; c <= uint32_t(kApplyClasses); ++c) cur.cg[c] = g0;
                cur.cg[kApplyClasses] = g1;
            }
            for (uint32_t g = g0; g < g1; ++g) {
                if (out.grp[g].u >= 0) out.ug[cur.u0 + uint32_t(out.grp[g].u)] = g;
                if (out.grp[g].m & far_mask) ++out.nfar;
            }
            out.chunk.push_back(cur);
            out.max_p = std::max(out.max_p, cur.p1 - cur.p0);
            out.max_u = std::max(out.max_u, cur.u1 - cur.u0);
        }
        cur = ApplyChunk{};
        cur.cg[0] = g1;
        cur.p0 = uint32_t(out.pzl.size());
        cur.u0 = uint32_t(out.ug.size());
    };
    auto cls = [tmask](uint64_t zl) {
        if ((zl & ~tmask) == 0) return 0;  // thread bits only (or zl == 0)
        if ((zl & tmask) == 0) return 2;   // row-index bits only
        return 1;                          // straddling
    };
    for (uint64_t g = 0; g < op->G; ++g) {
        std::vector<uint64_t> zls;
        std::vector<std::vector<uint32_t>> members;
        for (uint32_t i = op->gterm[g]; i < op->gterm[g + 1]; ++i) {
            const uint64_t zl = pairs ? pair_row_bits(op->tz[i] & mask, tb) : op->tz[i] & mask;
            size_t k = 0;
            while (k < zls.size() && zls[k] != zl) ++k;
            if (k == zls.size()) { zls.push_back(zl); members.emplace_back(); }
            members[k].push_back(i);
        }
        std::vector<size_t> order;  // patterns by class, first-seen order inside a class
        for (int c = 0; c < 3; ++c)
            for (size_t k = 0; k < zls.size(); ++k)
                if (cls(zls[k]) == c) order.push_back(k);
        for (size_t s0 = 0; s0 < order.size(); s0 += pcap) {
            const size_t s1 = std::min(order.size(), s0 + pcap);
            uint32_t ncls[3] = {0, 0, 0};
            for (size_t s = s0; s < s1; ++s) ++ncls[cls(zls[order[s]])];
            const uint32_t n = uint32_t(s1 - s0), has_u = ncls[2] ? 1u : 0u;
            if ((uint32_t(out.pzl.size()) - cur.p0) + n > pcap ||
                (uint32_t(out.ug.size()) - cur.u0) + has_u > ucap)
                close_chunk();
            ApplyGroup a{};
            a.m = op->gx[g];
            a.p0 = uint32_t(out.pzl.size()) - cur.p0;
            a.ps = a.p0 + ncls[0];
            a.pk = a.ps + ncls[1];
            a.p1 = a.pk + ncls[2];
            bool real = true;
            for (size_t s = s0; s < s1; ++s) {
                const size_t k = order[s];
                out.pzl.push_back(uint32_t(zls[k]));
                for (uint32_t i : members[k]) {
                    out.ptz.push_back(op->tz[i]);
                    out.ptc.push_back(op->tc[2 * i]);
                    out.ptc.push_back(op->tc[2 * i + 1]);
                    real = real && op->tc[2 * i + 1] == 0.0;
                }
                out.ptptr.push_back(uint32_t(out.ptz.size()));
            }
            a.u = has_u ? int32_t(uint32_t(out.ug.size()) - cur.u0) : -1;
            if (has_u) out.ug.push_back(uint32_t(out.grp.size()));
            a.flags = real ? kGroupReal : 0u;
            out.grp.push_back(a);
        }
    }
    close_chunk();
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src) {
    size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
    QR_HIP_CHECK(hipMalloc((void **)dst, bytes));
    if (!src.empty()) QR_HIP_CHECK(hipMemcpy(*dst, src.data(), src.size() * sizeof(T),
                                             hipMemcpyHostToDevice));
    return QR_OK;
}

int upload_tail(const qr_op *op, DeviceTables &t) {
    if (!op->ntail) return QR_OK;
    std::vector<double2> c2(op->ttz.size());
    for (size_t i = 0; i < c2.size(); ++i) c2[i] = make_double2(op->ttc[2 * i], op->ttc[2 * i + 1]);
    int rc;
    if ((rc = upload(&t.ttz, op->ttz)) || (rc = upload(&t.ttc, c2))) return rc;
    t.ntail = op->ntail;
    return QR_OK;
}

int device_tables(const qr_op *op, DeviceTables **out, bool noalloc) {
    int dev = 0;
    QR_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(op->mu);
    DeviceTables &t = op->dev[dev];
    if (!t.ready && noalloc)
        return fail(QR_ERR_ARG, "apply_noalloc: the operator's tables are not on this device "
                                "(one qr_pauli_apply of it on this device uploads them)");
    if (!t.ready) {
        int rc;
        std::vector<double2> tc2(op->T);
        for (uint64_t i = 0; i < op->T; ++i) tc2[i] = make_double2(op->tc[2 * i], op->tc[2 * i + 1]);
        if ((rc = upload(&t.gx, op->gx)) || (rc = upload(&t.gterm, op->gterm)) ||
            (rc = upload(&t.tz, op->tz)) || (rc = upload(&t.tc, tc2)) ||
            (rc = upload(&t.rl_ptr, op->rl_ptr)) || (rc = upload(&t.rl, op->rl)) ||
            (rc = upload(&t.gkind, op->gkind)) ||
            (!op->heavy.empty() && (rc = upload(&t.heavy, op->heavy))) ||
            (!op->leaf_heavy.empty() && (rc = upload(&t.leaf_heavy, op->leaf_heavy))) ||
            (rc = upload_tail(op, t))) {
            free_tables(t);
            return rc;
        }
        t.nheavy = uint32_t(op->heavy.size());
        t.ready = true;
    }
    *out = &t;
    return QR_OK;
}

bool tables_ready(const qr_op *op, const DeviceTables *t, int slot, int L) {
    std::lock_guard<std::mutex> lock(op->mu);
    return t->pat[slot].L == L;
}

int device_patterns(const qr_op *op, DeviceTables *t, int slot, int L, int tb, uint32_t pcap,
                    uint32_t ucap, bool staged, bool pairs, bool noalloc) {
    std::lock_guard<std::mutex> lock(op->mu);
    DeviceTables::Patterns &pt = t->pat[slot];
    if (pt.L == L) return QR_OK;
    if (noalloc)
        return fail(QR_ERR_ARG, "apply_noalloc: the operator's apply tables for this precision "
                                "are not on this device (one qr_pauli_apply of it with this "
                                "precision on this device uploads them)");
    for (void *p : {(void *)pt.grp, (void *)pt.chunk, (void *)pt.run, (void *)pt.ug, (void *)pt.pzl,
                    (void *)pt.ptptr, (void *)pt.ptz, (void *)pt.ptc})
        if (p) (void)hipFree(p);
    pt = DeviceTables::Patterns{};
    HostPatterns hp;
    build_patterns(op, L, tb, pcap, ucap, staged, pairs, hp);
    std::vector<double2> c2(hp.ptz.size());
    for (size_t i = 0; i < c2.size(); ++i) c2[i] = make_double2(hp.ptc[2 * i], hp.ptc[2 * i + 1]);
    int rc;
    if ((rc = upload(&pt.grp, hp.grp)) || (rc = upload(&pt.chunk, hp.chunk)) ||
        (staged && (rc = upload(&pt.run, hp.run))) ||
        (rc = upload(&pt.ug, hp.ug)) || (rc = upload(&pt.pzl, hp.pzl)) ||
        (rc = upload(&pt.ptptr, hp.ptptr)) || (rc = upload(&pt.ptz, hp.ptz)) ||
        (rc = upload(&pt.ptc, c2)))
        return rc;
    pt.P = uint32_t(hp.pzl.size());
    pt.ngrp = uint32_t(hp.grp.size());
    pt.nchunk = uint32_t(hp.chunk.size());
    pt.nrun = uint32_t(hp.run.size());
    pt.nfar = hp.nfar;
    pt.max_p = hp.max_p;
    pt.max_u = hp.max_u;
    pt.L = L;
    return QR_OK;
}

// The parts of an operator derived from its grouped tables (gx, gterm):
// the sub-operators (subtrees of at most kSubMaxG groups) and heavy groups.
void finish_op(qr_op *op) {
    // subtrees of at most kSubMaxG groups, split where build_rank_lists splits
    if (op->G > kSubMaxG) {
        struct Node { uint32_t lo, hi, depth; };
        std::vector<Node> stack{{0, uint32_t(op->G), 0}};
        while (!stack.empty()) {
            const Node nd = stack.back();
            stack.pop_back();
            if (nd.hi - nd.lo <= kSubMaxG) {
                op->leaves.push_back({nd.lo, nd.hi, nd.depth, 0, 0});
                continue;
            }
            const uint64_t bit = uint64_t(1) << (63 - __builtin_clzll(op->gx[nd.lo] ^ op->gx[nd.hi - 1]));
            const uint32_t m = uint32_t(std::partition_point(op->gx.begin() + nd.lo,
                                                             op->gx.begin() + nd.hi,
                                                             [bit](uint64_t v) { return (v & bit) == 0; }) -
                                        op->gx.begin());
            stack.push_back({m, nd.hi, nd.depth + 1});
            stack.push_back({nd.lo, m, nd.depth + 1});
        }
    }
    // heavy groups: at least 8 terms and 4x the mean, the largest kMaxHeavy
    {
        const double mean = double(op->T) / double(op->G);
        std::vector<std::pair<uint32_t, uint32_t>> cand;  // (terms, group)
        for (uint64_t g = 0; g < op->G; ++g) {
            const uint32_t n = op->gterm[g + 1] - op->gterm[g];
            if (n >= 8 && double(n) >= 4.0 * mean) cand.emplace_back(n, uint32_t(g));
        }
        std::sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) {
            return a.first != b.first ? a.first > b.first : a.second < b.second;
        });
        if (cand.size() > kMaxHeavy) cand.resize(kMaxHeavy);
        for (const auto &c : cand) op->heavy.push_back(c.second);
        std::sort(op->heavy.begin(), op->heavy.end());
        // each sub-operator's heavy groups, in its own (leaf-local) indices
        for (auto &lf : op->leaves) {
            lf.hoff = uint32_t(op->leaf_heavy.size());
            for (uint32_t g : op->heavy)
                if (g >= lf.s0 && g < lf.s1) op->leaf_heavy.push_back(g - lf.s0);
            lf.hcnt = uint32_t(op->leaf_heavy.size()) - lf.hoff;
        }
    }
    // tail table: operators with heavy groups, whose other groups also run
    // past two terms (the molecular Hamiltonians: 4 and 6)
    if (!op->heavy.empty()) {
        uint32_t nt = 0;
        for (uint64_t g = 0; g < op->G; ++g) {
            const uint32_t n = op->gterm[g + 1] - op->gterm[g];
            if (n > 2 && !std::binary_search(op->heavy.begin(), op->heavy.end(), uint32_t(g)))
                nt = std::max(nt, n - 2);
        }
        nt = std::min(nt, kMaxTail);
        while (nt && uint64_t(nt) * op->G > kMaxTailSlots) --nt;
        op->ntail = nt;
        op->ttz.assign(uint64_t(nt) * op->G, 0);
        op->ttc.assign(2 * uint64_t(nt) * op->G, 0.0);
        for (uint64_t g = 0; g < op->G; ++g) {
            const uint32_t t0 = op->gterm[g], n = op->gterm[g + 1] - t0;
            for (uint32_t u = 0; u < nt && u + 2 < n; ++u) {
                const uint64_t slot = uint64_t(u) * op->G + g;
                op->ttz[slot] = op->tz[t0 + 2 + u];
                op->ttc[2 * slot] = op->tc[2 * (t0 + 2 + u)];
                op->ttc[2 * slot + 1] = op->tc[2 * (t0 + 2 + u) + 1];
            }
        }
    }
}

}  // namespace qr

using namespace qr;

extern "C" {

const char *qr_last_error(void) { return g_last_error.c_str(); }

int qr_abi_version(void) { return QRHIP_ABI_VERSION; }

int qr_device_alloc(uint64_t bytes, void **d_out) {
    if (!d_out) return fail(QR_ERR_ARG, "null argument");
    *d_out = nullptr;
    if (bytes == 0) return QR_OK;
    const hipError_t e = hipMalloc(d_out, bytes);
    if (e == hipErrorOutOfMemory) return fail(QR_ERR_NOMEM, "device out of memory");
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    return QR_OK;
}

int qr_device_free(void *d_ptr) {
    if (d_ptr) QR_HIP_CHECK(hipFree(d_ptr));
    return QR_OK;
}

int qr_copy_to_device(void *d_dst, const void *src, uint64_t bytes, void *stream) {
    if (bytes == 0) return QR_OK;
    if (!d_dst || !src) return fail(QR_ERR_ARG, "null argument");
    QR_HIP_CHECK(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return QR_OK;
}

int qr_copy_to_host(void *dst, const void *d_src, uint64_t bytes, void *stream) {
    if (bytes == 0) return QR_OK;
    if (!dst || !d_src) return fail(QR_ERR_ARG, "null argument");
    QR_HIP_CHECK(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    QR_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return QR_OK;
}

int qr_pauli_parse(const char *label, uint32_t *num_qubits, uint32_t *base_phase,
                   uint32_t *phase, uint64_t *x_mask, uint64_t *z_mask) {
    uint32_t nq, bp, ny;
    uint64_t x, z;
    int rc = parse_label(label, &nq, &bp, &ny, &x, &z);
    if (rc) return rc;
    if (num_qubits) *num_qubits = nq;
    if (base_phase) *base_phase = bp;
    if (phase) *phase = (bp + ny) % 4;
    if (x_mask) *x_mask = x;
    if (z_mask) *z_mask = z;
    return QR_OK;
}

int qr_op_create(const char *const *labels, const double *coeffs, uint64_t num_terms,
                 int32_t phase_convention, qr_op **out) {
#pragma clang fp contract(off)
    if (!out) return fail(QR_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (num_terms == 0)
        return fail(QR_ERR_TERMS, "SparsePauliOp::new: at least one pauli must be supplied");
    if (!labels || !coeffs) return fail(QR_ERR_ARG, "null labels or coefficients");
    if (phase_convention != QR_PHASE_ROWWISE && phase_convention != QR_PHASE_KRON)
        return fail(QR_ERR_ARG, "unknown phase convention");
    if (num_terms > 0xFFFFFFF0ull) return fail(QR_ERR_SIZE, "too many terms");
    qr_op *op = new (std::nothrow) qr_op;
    if (!op) return fail(QR_ERR_NOMEM, "out of host memory");
    const uint64_t T = num_terms;
    op->T = T;
    op->x.resize(T);
    op->z.resize(T);
    op->cp.resize(2 * T);
    // M[p] of accel.rs:147-152, multiplied as num-complex does:
    // (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) with a = M[p], b = c.
    static const double M[4][2] = {{1.0, 0.0}, {0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}};
    for (uint64_t t = 0; t < T; ++t) {
        uint32_t nq, bp, ny;
        int rc = parse_label(labels[t], &nq, &bp, &ny, &op->x[t], &op->z[t]);
        if (rc) { delete op; return rc; }
        if (t == 0) op->nq = nq;
        else if (nq != op->nq) {
            delete op;
            return fail(QR_ERR_TERMS,
                        "SparsePauliOp::new: all supplied paulis must have the same #qubits");
        }
        uint32_t p = phase_convention == QR_PHASE_ROWWISE ? (bp + ny) % 4 : (ny + 4 - bp) % 4;
        op->cp[2 * t] = M[p][0] * coeffs[2 * t] - M[p][1] * coeffs[2 * t + 1];
        op->cp[2 * t + 1] = M[p][0] * coeffs[2 * t + 1] + M[p][1] * coeffs[2 * t];
    }
    // group by x: stable sort keeps input order inside a group
    std::vector<uint32_t> order(T);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [op](uint32_t a, uint32_t b) { return op->x[a] < op->x[b]; });
    op->tz.resize(T);
    op->tc.resize(2 * T);
    for (uint64_t i = 0; i < T; ++i) {
        uint32_t t = order[i];
        if (i == 0 || op->x[t] != op->gx.back()) {
            op->gx.push_back(op->x[t]);
            op->gterm.push_back(uint32_t(i));
        }
        op->tz[i] = op->z[t];
        op->tc[2 * i] = op->cp[2 * t];
        op->tc[2 * i + 1] = op->cp[2 * t + 1];
    }
    op->gterm.push_back(uint32_t(T));
    op->G = op->gx.size();
    op->gkind.resize(op->G);
    for (uint64_t g = 0; g < op->G; ++g) {
        bool real = true, imag = true;
        for (uint32_t i = op->gterm[g]; i < op->gterm[g + 1]; ++i) {
            if (op->tc[2 * i + 1] != 0.0) real = false;
            if (op->tc[2 * i] != 0.0) imag = false;
        }
        op->gkind[g] = real ? 1 : (imag ? 2 : 0);
    }
    build_rank_lists(op->gx, op->rl_ptr, op->rl);
    finish_op(op);
    *out = op;
    return QR_OK;
}

int qr_op_create_device(const char *const *labels, const double *coeffs, uint64_t num_terms,
                        int32_t phase_convention, void *stream, qr_op **out) {
    if (!out) return fail(QR_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (num_terms == 0)
        return fail(QR_ERR_TERMS, "SparsePauliOp::new: at least one pauli must be supplied");
    if (!labels || !coeffs) return fail(QR_ERR_ARG, "null labels or coefficients");
    if (phase_convention != QR_PHASE_ROWWISE && phase_convention != QR_PHASE_KRON)
        return fail(QR_ERR_ARG, "unknown phase convention");
    if (num_terms > 0xFFFFFFF0ull) return fail(QR_ERR_SIZE, "too many terms");
    return group_terms_device(labels, coeffs, num_terms, phase_convention, (hipStream_t)stream, out);
}

int qr_op_tables(const qr_op *op, uint64_t *gx, uint32_t *gterm, uint64_t *tz, double *tc,
                 uint8_t *gkind, uint32_t *rl_ptr, int32_t *rl, uint64_t *num_rank_entries) {
    if (!op) return fail(QR_ERR_ARG, "null operator");
    if (gx) std::copy(op->gx.begin(), op->gx.end(), gx);
    if (gterm) std::copy(op->gterm.begin(), op->gterm.end(), gterm);
    if (tz) std::copy(op->tz.begin(), op->tz.end(), tz);
    if (tc) std::copy(op->tc.begin(), op->tc.end(), tc);
    if (gkind) std::copy(op->gkind.begin(), op->gkind.end(), gkind);
    if (rl_ptr) std::copy(op->rl_ptr.begin(), op->rl_ptr.end(), rl_ptr);
    if (rl)
        for (size_t e = 0; e < op->rl.size(); ++e) {
            rl[2 * e] = op->rl[e].bit;
            rl[2 * e + 1] = op->rl[e].sigma;
        }
    if (num_rank_entries) *num_rank_entries = op->rl.size();
    return QR_OK;
}

int qr_op_destroy(qr_op *op) {
    if (!op) return QR_OK;
    int cur = 0;
    bool have = hipGetDevice(&cur) == hipSuccess;
    for (int d = 0; d < kMaxDevices; ++d) {
        if (!op->dev[d].ready && !op->dev[d].gx) continue;
        if (hipSetDevice(d) == hipSuccess) free_tables(op->dev[d]);
    }
    if (have) (void)hipSetDevice(cur);
    delete op;
    return QR_OK;
}

int qr_op_info(const qr_op *op, uint32_t *num_qubits, uint64_t *num_terms, uint64_t *num_groups,
               uint64_t *nnz) {
    if (!op) return fail(QR_ERR_ARG, "null operator");
    if (num_qubits) *num_qubits = op->nq;
    if (num_terms) *num_terms = op->T;
    if (num_groups) *num_groups = op->G;
    if (nnz) {
        unsigned __int128 v = (unsigned __int128)op->G << op->nq;
        *nnz = (v >> 64) ? 0 : uint64_t(v);
    }
    return QR_OK;
}

int qr_op_terms(const qr_op *op, uint64_t *x, uint64_t *z, double *coeff) {
    if (!op) return fail(QR_ERR_ARG, "null operator");
    if (x) std::memcpy(x, op->x.data(), op->T * sizeof(uint64_t));
    if (z) std::memcpy(z, op->z.data(), op->T * sizeof(uint64_t));
    if (coeff) std::memcpy(coeff, op->cp.data(), 2 * op->T * sizeof(double));
    return QR_OK;
}

int qr_op_row_layout(const qr_op *op, uint64_t row, uint64_t *cols, uint32_t *groups) {
    if (!op) return fail(QR_ERR_ARG, "null operator");
    if (op->nq < 64 && (row >> op->nq) != 0) return fail(QR_ERR_ARG, "row out of range");
    for (uint64_t j = 0; j < op->G; ++j) {
        int64_t rank = int64_t(j);
        for (uint32_t e = op->rl_ptr[j]; e < op->rl_ptr[j + 1]; ++e)
            if ((row >> op->rl[e].bit) & 1) rank += op->rl[e].sigma;
        if (rank < 0 || uint64_t(rank) >= op->G)
            return fail(QR_ERR_ARG, "internal error: rank out of range");
        if (cols) cols[rank] = row ^ op->gx[j];
        if (groups) groups[rank] = uint32_t(j);
    }
    return QR_OK;
}

int qr_mode_parse(const char *mode, int32_t *kind, uint64_t *step, int32_t *phase_convention) {
    if (!mode) return fail(QR_ERR_ARG, "null mode");
    std::string m(mode);
    int32_t k = -1;
    uint64_t st = 0;
    auto find_step = [&](const char *key) -> int {
        // unanchored search like Regex::is_match(r"<key>/(\d+)") (lib.rs:283-288)
        size_t klen = std::strlen(key);
        for (size_t pos = m.find(key); pos != std::string::npos; pos = m.find(key, pos + 1)) {
            size_t d = pos + klen;
            if (d < m.size() && m[d] == '/' && d + 1 < m.size() && m[d + 1] >= '0' && m[d + 1] <= '9') {
                uint64_t v = 0;
                bool overflow = false;
                for (size_t i = d + 1; i < m.size() && m[i] >= '0' && m[i] <= '9'; ++i) {
                    if (v > (UINT64_MAX - 9) / 10) overflow = true;
                    v = v * 10 + uint64_t(m[i] - '0');
                }
                st = overflow ? 0 : v;
                return 1;
            }
        }
        return 0;
    };
    if (m.empty()) k = QR_MODE_DEFAULT;
    else if (m == "Binary") k = QR_MODE_BINARY;
    else if (m == "Accel") k = QR_MODE_ACCEL;
    else if (m == "Rowwise") k = QR_MODE_ROWWISE;
    else if (m == "RowwiseUnsafe") k = QR_MODE_ROWWISE_UNSAFE;
    else if (find_step("RowwiseUnsafeChunked")) k = QR_MODE_ROWWISE_UNSAFE_CHUNKED;
    else if (m == "Reduce") k = QR_MODE_REDUCE;
    else if (m == "Rayon") k = QR_MODE_RAYON;
    else if (find_step("RayonChunked")) k = QR_MODE_RAYON_CHUNKED;
    else if (m == "Hip") k = QR_MODE_HIP;
    if (k < 0) return fail(QR_ERR_MODE, "to_matrix_mode: unrecognized mode " + m);
    if ((k == QR_MODE_ROWWISE_UNSAFE_CHUNKED || k == QR_MODE_RAYON_CHUNKED) && st == 0)
        return fail(QR_ERR_MODE, "to_matrix_mode: chunk size must be a positive integer in " + m);
    if (kind) *kind = k;
    if (step) *step = st;
    if (phase_convention) {
        bool kron = (k == QR_MODE_DEFAULT || k == QR_MODE_BINARY || k == QR_MODE_REDUCE ||
                     k == QR_MODE_RAYON);
        *phase_convention = kron ? QR_PHASE_KRON : QR_PHASE_ROWWISE;
    }
    return QR_OK;
}

int qr_csr_build(const qr_op *op, uint64_t *d_indptr, uint64_t *d_indices, double *d_data,
                 void *stream) {
    if (!op || !d_indptr || !d_indices || !d_data) return fail(QR_ERR_ARG, "null argument");
    if (op->nq > 40) return fail(QR_ERR_SIZE, "CSR of more than 40 qubits is not supported");
    DeviceTables *t;
    int rc = device_tables(op, &t);
    if (rc) return rc;
    return launch_csr_build(op, t, 0, uint64_t(1) << op->nq, d_indptr, d_indices, d_data,
                            (hipStream_t)stream);
}

int qr_csr_build_rows(const qr_op *op, uint64_t row_begin, uint64_t row_end, uint64_t *d_indices,
                      double *d_data, void *stream) {
    if (!op || !d_indices || !d_data) return fail(QR_ERR_ARG, "null argument");
    if (op->nq > 40) return fail(QR_ERR_SIZE, "CSR of more than 40 qubits is not supported");
    if (row_end < row_begin || row_end > (uint64_t(1) << op->nq))
        return fail(QR_ERR_ARG, "row range out of bounds");
    if (row_end == row_begin) return QR_OK;
    DeviceTables *t;
    int rc = device_tables(op, &t);
    if (rc) return rc;
    return launch_csr_build(op, t, row_begin, row_end, nullptr, d_indices, d_data,
                            (hipStream_t)stream);
}

int qr_csr_matvec(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                  const uint64_t *d_indices, const double *d_data, const double *d_x, double *d_y,
                  void *stream) {
    if (nrows == 0) return QR_OK;
    if (!d_indptr || !d_y) return fail(QR_ERR_ARG, "null argument");
    if (nnz != 0 && (!d_indices || !d_data || !d_x || ncols == 0))
        return fail(QR_ERR_ARG, "null argument");
    return launch_csr_matvec(nrows, ncols, nnz, d_indptr, d_indices, d_data, d_x, d_y,
                             (hipStream_t)stream);
}

int qr_csr_diagonal(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
                    const uint64_t *d_indices, const double *d_data, double *d_out, void *stream) {
    if (nrows == 0 || ncols == 0) return QR_OK;
    if (!d_indptr || !d_out || (nnz && (!d_indices || !d_data)))
        return fail(QR_ERR_ARG, "null argument");
    return launch_csr_diagonal(nrows, ncols, nnz, d_indptr, d_indices, d_data, d_out,
                               (hipStream_t)stream);
}

int qr_axpby(uint64_t n, const double *a, const double *d_x, const double *b, const double *d_y,
             double *d_z, void *stream) {
    if (!a || !b || (n && (!d_x || !d_y || !d_z))) return fail(QR_ERR_ARG, "null argument");
    return launch_axpby(n, 0, a, d_x, b, d_y, d_z, (hipStream_t)stream);
}

int qr_axpy(uint64_t n, const double *a, const double *d_x, const double *d_y, double *d_z,
            void *stream) {
    if (!a || (n && (!d_x || !d_y || !d_z))) return fail(QR_ERR_ARG, "null argument");
    return launch_axpby(n, 1, a, d_x, nullptr, d_y, d_z, (hipStream_t)stream);
}

int qr_ax(uint64_t n, const double *a, const double *d_x, double *d_z, void *stream) {
    if (!a || (n && (!d_x || !d_z))) return fail(QR_ERR_ARG, "null argument");
    return launch_axpby(n, 2, a, d_x, nullptr, nullptr, d_z, (hipStream_t)stream);
}

int qr_precond(uint64_t nrows, uint64_t ncols, uint64_t nnz, const uint64_t *d_indptr,
               const uint64_t *d_indices, const double *d_data, const double *d_dx,
               const double *e, double tol, double *d_out, void *stream) {
    const uint64_t n = nrows < ncols ? nrows : ncols;
    if (!e || (n && (!d_indptr || !d_dx || !d_out || (nnz && (!d_indices || !d_data)))))
        return fail(QR_ERR_ARG, "null argument");
    return launch_precond(n, nullptr, nnz, d_indptr, d_indices, d_data, d_dx, e, tol, d_out,
                          (hipStream_t)stream);
}

int qr_precond2(uint64_t n, const double *d_diag, const double *d_dx, const double *e, double tol,
                double *d_out, void *stream) {
    if (!e || (n && (!d_diag || !d_dx || !d_out))) return fail(QR_ERR_ARG, "null argument");
    return launch_precond(n, d_diag, 0, nullptr, nullptr, nullptr, d_dx, e, tol, d_out,
                          (hipStream_t)stream);
}

int qr_csr_add_count(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *d_indptr1,
                     const uint64_t *d_indices1, const double *d_data1, const double *b,
                     uint64_t nnz2, const uint64_t *d_indptr2, const uint64_t *d_indices2,
                     const double *d_data2, uint64_t *d_indptr_out, uint64_t *nnz_out,
                     void *stream) {
    if (!a || !b || !nnz_out || !d_indptr_out) return fail(QR_ERR_ARG, "null argument");
    *nnz_out = 0;
    if (nrows == 0) {
        QR_HIP_CHECK(hipMemsetAsync(d_indptr_out, 0, sizeof(uint64_t), (hipStream_t)stream));
        return QR_OK;
    }
    if (!d_indptr1 || !d_indptr2 || (nnz1 && (!d_indices1 || !d_data1)) ||
        (nnz2 && (!d_indices2 || !d_data2)))
        return fail(QR_ERR_ARG, "null argument");
    return launch_spadd_count(nrows, a, nnz1, d_indptr1, d_indices1, d_data1, b, nnz2, d_indptr2,
                              d_indices2, d_data2, d_indptr_out, nnz_out, (hipStream_t)stream);
}

int qr_csr_add_fill(uint64_t nrows, const double *a, uint64_t nnz1, const uint64_t *d_indptr1,
                    const uint64_t *d_indices1, const double *d_data1, const double *b,
                    uint64_t nnz2, const uint64_t *d_indptr2, const uint64_t *d_indices2,
                    const double *d_data2, const uint64_t *d_indptr_out, uint64_t *d_indices_out,
                    double *d_data_out, void *stream) {
    if (!a || !b || !d_indptr_out) return fail(QR_ERR_ARG, "null argument");
    if (nrows == 0) return QR_OK;
    if (!d_indptr1 || !d_indptr2 || (nnz1 && (!d_indices1 || !d_data1)) ||
        (nnz2 && (!d_indices2 || !d_data2)))
        return fail(QR_ERR_ARG, "null argument");
    return launch_spadd_fill(nrows, a, nnz1, d_indptr1, d_indices1, d_data1, b, nnz2, d_indptr2,
                             d_indices2, d_data2, d_indptr_out, d_indices_out, d_data_out,
                             (hipStream_t)stream);
}

int qr_csr_count_zeros(uint64_t nnz, const double *d_data, double tol, uint64_t *count,
                       void *stream) {
    if (!count || (nnz && !d_data)) return fail(QR_ERR_ARG, "null argument");
    return launch_count_zeros(nnz, d_data, tol, count, (hipStream_t)stream);
}

int qr_csr_eliminate_zeros(uint64_t nrows, uint64_t nnz, const uint64_t *d_indptr,
                           const uint64_t *d_indices, const double *d_data, double tol,
                           uint64_t *d_indptr_out, uint64_t *d_indices_out, double *d_data_out,
                           uint64_t *nnz_out, void *stream) {
    if (!nnz_out || !d_indptr || !d_indptr_out ||
        (nnz && (!d_indices || !d_data || !d_indices_out || !d_data_out)))
        return fail(QR_ERR_ARG, "null argument");
    *nnz_out = 0;
    return launch_eliminate_zeros(nrows, nnz, d_indptr, d_indices, d_data, tol, d_indptr_out,
                                  d_indices_out, d_data_out, nnz_out, (hipStream_t)stream);
}

int qr_csr_write_raw(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz,
                     const uint64_t *d_indptr, const uint64_t *d_indices, const double *d_data,
                     void *stream) {
    if (!path || !d_indptr || (nnz && (!d_indices || !d_data))) return fail(QR_ERR_ARG, "null argument");
    return io_write_raw(path, rows, cols, nnz, d_indptr, d_indices, d_data, (hipStream_t)stream);
}

int qr_csr_read_raw_info(const char *path, uint64_t *rows, uint64_t *cols, uint64_t *nnz) {
    if (!path || !rows || !cols || !nnz) return fail(QR_ERR_ARG, "null argument");
    return io_read_raw_info(path, rows, cols, nnz);
}

int qr_csr_read_raw(const char *path, uint64_t rows, uint64_t nnz, uint64_t *d_indptr,
                    uint64_t *d_indices, double *d_data, void *stream) {
    if (!path || !d_indptr || (nnz && (!d_indices || !d_data))) return fail(QR_ERR_ARG, "null argument");
    return io_read_raw(path, rows, nnz, d_indptr, d_indices, d_data, (hipStream_t)stream);
}

int qr_csr_write_mm(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz,
                    const uint64_t *d_indptr, const uint64_t *d_indices, const double *d_data,
                    int32_t symmetry, void *stream) {
    if (!path || !d_indptr || (nnz && (!d_indices || !d_data))) return fail(QR_ERR_ARG, "null argument");
    return io_write_mm(path, rows, cols, nnz, d_indptr, d_indices, d_data, symmetry,
                       (hipStream_t)stream);
}

int qr_mm_read_open(const char *path, qr_mm_reader **out, uint64_t *rows, uint64_t *cols,
                    uint64_t *nnz, void *stream) {
    if (!path || !out || !rows || !cols || !nnz) return fail(QR_ERR_ARG, "null argument");
    *out = nullptr;
    const int rc = io_read_mm(path, out, (hipStream_t)stream);
    if (rc != QR_OK) return rc;
    *rows = (*out)->rows;
    *cols = (*out)->cols;
    *nnz = (*out)->nnz;
    return QR_OK;
}

int qr_mm_read_take(const qr_mm_reader *r, uint64_t *d_indptr, uint64_t *d_indices, double *d_data,
                    void *stream) {
    if (!r || !d_indptr || (r->nnz && (!d_indices || !d_data))) return fail(QR_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    QR_HIP_CHECK(hipMemcpyAsync(d_indptr, r->indptr, (r->rows + 1) * 8, hipMemcpyDeviceToDevice, s));
    if (r->nnz) {
        QR_HIP_CHECK(hipMemcpyAsync(d_indices, r->indices, r->nnz * 8, hipMemcpyDeviceToDevice, s));
        QR_HIP_CHECK(hipMemcpyAsync(d_data, r->data, r->nnz * 16, hipMemcpyDeviceToDevice, s));
    }
    QR_HIP_CHECK(hipStreamSynchronize(s));
    return QR_OK;
}

int qr_mm_read_close(qr_mm_reader *r) {
    delete r;
    return QR_OK;
}

static int apply_args(const qr_op *op, int32_t precision, uint32_t nvec, const void *d_x,
                      void *d_y, double *d_expect) {
    if (!op || !d_x || (!d_y && !d_expect)) return fail(QR_ERR_ARG, "null argument");
    if (precision != QR_PREC_FP64 && precision != QR_PREC_MIXED)
        return fail(QR_ERR_ARG, "unknown precision");
    if (nvec == 0) return QR_OK;
    if (op->nq > 36) return fail(QR_ERR_SIZE, "states of more than 36 qubits are not supported");
    if (d_y) {
        // y is written while other blocks still gather x: no overlap allowed
        const uint64_t bytes = uint64_t(nvec) << (op->nq + (precision == QR_PREC_MIXED ? 3 : 4));
        const uintptr_t x0 = uintptr_t(d_x), y0 = uintptr_t(d_y);
        if (x0 < y0 + bytes && y0 < x0 + bytes)
            return fail(QR_ERR_ARG, "apply: y overlaps x (the apply is not in place)");
    }
    return QR_OK;
}

int qr_pauli_apply(const qr_op *op, int32_t precision, uint32_t nvec, const void *d_x, void *d_y,
                   double *d_expect, void *stream) {
    int rc = apply_args(op, precision, nvec, d_x, d_y, d_expect);
    if (rc || nvec == 0) return rc;
    DeviceTables *t;
    if ((rc = device_tables(op, &t))) return rc;
    return launch_pauli_apply(op, t, precision, nvec, d_x, d_y, d_expect, (hipStream_t)stream);
}

int qr_pauli_apply_noalloc(const qr_op *op, int32_t precision, uint32_t nvec, const void *d_x,
                           void *d_y, double *d_expect, void *stream) {
    int rc = apply_args(op, precision, nvec, d_x, d_y, d_expect);
    if (rc || nvec == 0) return rc;
    DeviceTables *t;
    if ((rc = device_tables(op, &t, /*noalloc=*/true))) return rc;
    return launch_pauli_apply(op, t, precision, nvec, d_x, d_y, d_expect, (hipStream_t)stream,
                              /*noalloc=*/true);
}

}  // extern "C"
