// qr_io.hip -- files of a device-resident CSR matrix (complex128 values,
// uint64 indptr and indices):
//
//  * the raw binary format of qrusty::rawio (qrusty/src/rawio.rs:128-179):
//      "MI" (2 bytes; read back as "IM" means the other byte order),
//      u64 storage (0 = CSR), u64 rows, u64 cols,
//      u64 len + indptr[len], u64 len + indices[len], u64 len + data[len]
//      (data as re, im f64 pairs), all in the writer's byte order;
//  * MatrixMarket coordinate complex files (the sprs 0.11 io module, as
//    pyqrusty calls it, pyqrusty/src/lib.rs:216-259; the general writer is
//    restated in the reference as util::write_matrix_market_for_scipy,
//    qrusty/src/util.rs:172-199): a header line, "% written by sprs", the
//    "rows cols nnz" line, then "row col re im" per entry, 1-based, with
//    Rust's Display for the floats.
//
// A path whose extension is "gz" is gzip-compressed, as the reference's
// with_input_file / with_output_file decide (qrusty/src/util.rs:316-369).
//
// The matrix never makes a whole-size host copy on the way out: entries are
// streamed from the device through two pinned buffers (the copy of the next
// chunk runs while the current one is formatted and written).  MatrixMarket
// text is formatted by up to 16 threads, and gzip output is deflated in
// parallel 1 MiB blocks that form one gzip member (each block primed with the
// 32 KiB before it and closed with a sync flush, CRC-32s combined in order).
// Reading MatrixMarket parses on the host threads, then sorts the triplets on
// the device (radix sort on row << 32 | col, stable, so duplicates keep file
// order), sums duplicates in that order and builds indptr with a binary
// search per row.
#include "qr_internal.h"

#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>
#include <zlib.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace qr {
namespace {

unsigned io_threads() {
    const unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 1u : std::min(n, 16u);
}

// f(k) for k in [0, n), one thread each (the caller's thread runs k = 0)
template <typename F>
void parallel_for(unsigned n, F f) {
    std::vector<std::thread> th;
    th.reserve(n);
    for (unsigned k = 1; k < n; ++k) th.emplace_back(f, k);
    if (n) f(0u);
    for (auto &t : th) t.join();
}

// Rust's Path::extension() == Some("gz"): the text after the last '.' of the
// final component, when that '.' is not the component's first character.
bool is_gz_path(const char *path) {
    const char *name = strrchr(path, '/');
    name = name ? name + 1 : path;
    const char *dot = strrchr(name, '.');
    return dot && dot != name && strcmp(dot + 1, "gz") == 0;
}

// ---------------------------------------------------------------- sinks ---
class Sink {
  public:
    virtual ~Sink() = default;
    virtual bool write(const char *p, size_t n) = 0;
    virtual bool finish() = 0;  // flush and close
};

class FileSink final : public Sink {
    FILE *f_;

  public:
    explicit FileSink(FILE *f) : f_(f) { setvbuf(f_, nullptr, _IOFBF, 1 << 22); }
    ~FileSink() override {
        if (f_) fclose(f_);
    }
    bool write(const char *p, size_t n) override { return fwrite(p, 1, n, f_) == n; }
    bool finish() override {
        bool ok = fflush(f_) == 0;
        ok = fclose(f_) == 0 && ok;
        f_ = nullptr;
        return ok;
    }
};

// raw deflate of `in` (primed with `dict`) ended by a sync flush, or, with
// `last`, by the final block
bool deflate_block(const std::string &dict, const char *in, size_t n, bool last, std::string &out) {
    z_stream z{};
    if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK)
        return false;
    bool ok = true;
    if (!dict.empty())
        ok = deflateSetDictionary(&z, reinterpret_cast<const Bytef *>(dict.data()),
                                  uInt(dict.size())) == Z_OK;
    out.resize(deflateBound(&z, uLong(n)) + 64);
    z.next_in = reinterpret_cast<Bytef *>(const_cast<char *>(in));
    z.avail_in = uInt(n);
    z.next_out = reinterpret_cast<Bytef *>(&out[0]);
    z.avail_out = uInt(out.size());
    if (ok) {
        const int rc = deflate(&z, last ? Z_FINISH : Z_SYNC_FLUSH);
        ok = last ? rc == Z_STREAM_END : (rc == Z_OK && z.avail_in == 0 && z.avail_out > 0);
    }
    out.resize(out.size() - z.avail_out);
    deflateEnd(&z);
    return ok;
}

class GzSink final : public Sink {
    FileSink out_;
    std::string pending_;  // input not yet compressed
    std::string dict_;     // the last 32 KiB of compressed input
    uLong crc_ = crc32(0L, Z_NULL, 0);
    uint64_t total_ = 0;
    bool ok_ = true;
    static constexpr size_t kBlock = size_t(1) << 20;
    static constexpr size_t kWindow = 32768;

    bool compress(size_t nblocks_max, bool all) {
        const size_t full = pending_.size() / kBlock;
        const size_t nb = all ? (pending_.size() + kBlock - 1) / kBlock : std::min(full, nblocks_max);
        if (nb == 0) return true;
        const size_t used = std::min(pending_.size(), nb * kBlock);
        std::vector<std::string> outs(nb);
        std::vector<uLong> crcs(nb);
        std::vector<char> oks(nb, 1);
        const unsigned T = io_threads();
        parallel_for(std::min<size_t>(T, nb), [&](unsigned t) {
            for (size_t b = t; b < nb; b += T) {
                const size_t b0 = b * kBlock, n = std::min(kBlock, used - b0);
                std::string dict;
                if (b == 0) dict = dict_;
                else dict.assign(pending_, b0 - kWindow, kWindow);
                oks[b] = deflate_block(dict, pending_.data() + b0, n, false, outs[b]);
                crcs[b] = crc32(0L, reinterpret_cast<const Bytef *>(pending_.data() + b0), uInt(n));
            }
        });
        for (size_t b = 0; b < nb; ++b) {
            const size_t n = std::min(kBlock, used - b * kBlock);
            ok_ = ok_ && oks[b] && out_.write(outs[b].data(), outs[b].size());
            crc_ = crc32_combine(crc_, crcs[b], z_off_t(n));
            total_ += n;
        }
        std::string tail = dict_ + pending_.substr(used > kWindow ? used - kWindow : 0,
                                                   std::min(used, kWindow));
        dict_ = tail.size() > kWindow ? tail.substr(tail.size() - kWindow) : tail;
        pending_.erase(0, used);
        return ok_;
    }

  public:
    explicit GzSink(FILE *f) : out_(f) {
        static const unsigned char hdr[10] = {0x1f, 0x8b, 8, 0, 0, 0, 0, 0, 0, 0xff};
        ok_ = out_.write(reinterpret_cast<const char *>(hdr), 10);
    }
    bool write(const char *p, size_t n) override {
        pending_.append(p, n);
        const size_t batch = kBlock * io_threads();
        if (pending_.size() >= batch) compress(pending_.size() / kBlock, false);
        return ok_;
    }
    bool finish() override {
        compress(0, true);
        std::string fin;
        ok_ = ok_ && deflate_block(dict_, "", 0, true, fin) && out_.write(fin.data(), fin.size());
        unsigned char tr[8];
        for (int k = 0; k < 4; ++k) tr[k] = (unsigned char)(crc_ >> (8 * k));
        for (int k = 0; k < 4; ++k) tr[4 + k] = (unsigned char)(total_ >> (8 * k));
        ok_ = ok_ && out_.write(reinterpret_cast<const char *>(tr), 8);
        return out_.finish() && ok_;
    }
};

std::unique_ptr<Sink> open_sink(const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) return nullptr;
    if (is_gz_path(path)) return std::make_unique<GzSink>(f);
    return std::make_unique<FileSink>(f);
}

// -------------------------------------------------------------- sources ---
class Source {
  public:
    virtual ~Source() = default;
    // reads up to n bytes; returns the count (< n only at the end); sets bad
    virtual size_t read(void *p, size_t n) = 0;
    bool bad = false;
};

class FileSource final : public Source {
    FILE *f_;

  public:
    explicit FileSource(FILE *f) : f_(f) {}
    ~FileSource() override { fclose(f_); }
    size_t read(void *p, size_t n) override {
        const size_t got = fread(p, 1, n, f_);
        if (got < n && ferror(f_)) bad = true;
        return got;
    }
};

class GzSource final : public Source {
    gzFile g_;

  public:
    explicit GzSource(gzFile g) : g_(g) { gzbuffer(g_, 1 << 20); }
    ~GzSource() override { gzclose(g_); }
    size_t read(void *p, size_t n) override {
        size_t got = 0;
        while (got < n) {
            const unsigned want = unsigned(std::min<size_t>(n - got, 1u << 30));
            const int r = gzread(g_, static_cast<char *>(p) + got, want);
            if (r < 0) {
                bad = true;
                break;
            }
            if (r == 0) break;
            got += size_t(r);
        }
        return got;
    }
};

std::unique_ptr<Source> open_source(const char *path) {
    if (is_gz_path(path)) {
        gzFile g = gzopen(path, "rb");
        if (!g) return nullptr;
        return std::make_unique<GzSource>(g);
    }
    FILE *f = fopen(path, "rb");
    if (!f) return nullptr;
    return std::make_unique<FileSource>(f);
}

std::string os_reason() { return std::string(strerror(errno)); }

// ------------------------------------------------- pinned double buffers ---
struct Pinned {
    static constexpr size_t kCap = size_t(32) << 20;
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ok = false;
    Pinned() {
        ok = hipHostMalloc(reinterpret_cast<void **>(&buf[0]), kCap, hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&buf[1]), kCap, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
    }
    ~Pinned() {
        for (int k = 0; k < 2; ++k) {
            if (ev[k]) (void)hipEventSynchronize(ev[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            if (buf[k]) (void)hipHostFree(buf[k]);
        }
    }
};

// device bytes [src, src + bytes) handed to consume(host, n) in order, the
// copy of each next chunk overlapping consume() of the current one
template <typename F>
int stream_out(Pinned &pin, const void *src, size_t bytes, size_t chunk, hipStream_t s, F consume) {
    if (bytes == 0) return QR_OK;
    chunk = std::min(chunk, Pinned::kCap);
    const char *d = static_cast<const char *>(src);
    int cur = 0;
    QR_HIP_CHECK(hipMemcpyAsync(pin.buf[0], d, std::min(chunk, bytes), hipMemcpyDeviceToHost, s));
    QR_HIP_CHECK(hipEventRecord(pin.ev[0], s));
    for (size_t off = 0; off < bytes;) {
        const size_t n = std::min(chunk, bytes - off), next = off + n;
        if (next < bytes) {
            QR_HIP_CHECK(hipMemcpyAsync(pin.buf[cur ^ 1], d + next, std::min(chunk, bytes - next),
                                        hipMemcpyDeviceToHost, s));
            QR_HIP_CHECK(hipEventRecord(pin.ev[cur ^ 1], s));
        }
        QR_HIP_CHECK(hipEventSynchronize(pin.ev[cur]));
        const int rc = consume(pin.buf[cur], n);
        if (rc != QR_OK) {
            (void)hipStreamSynchronize(s);
            return rc;
        }
        off = next;
        cur ^= 1;
    }
    return QR_OK;
}

void swab8(char *p, size_t n8) {
    uint64_t *q = reinterpret_cast<uint64_t *>(p);
    for (size_t i = 0; i < n8; ++i) q[i] = __builtin_bswap64(q[i]);
}

// `bytes` from the source into device memory at dst (8-byte words swapped
// when `swab`); the read of each next chunk overlaps the copy of the current
int stream_in(Pinned &pin, Source &src, void *dst, size_t bytes, bool swab, hipStream_t s,
              const char *what) {
    char *d = static_cast<char *>(dst);
    int cur = 0;
    bool used[2] = {false, false};
    for (size_t off = 0; off < bytes;) {
        const size_t n = std::min(Pinned::kCap, bytes - off);
        if (used[cur]) QR_HIP_CHECK(hipEventSynchronize(pin.ev[cur]));
        if (src.read(pin.buf[cur], n) != n) {
            (void)hipStreamSynchronize(s);
            return fail(src.bad ? QR_ERR_IO : QR_ERR_FORMAT,
                        std::string("raw read: unexpected end of file in ") + what);
        }
        if (swab) swab8(pin.buf[cur], n / 8);
        QR_HIP_CHECK(hipMemcpyAsync(d + off, pin.buf[cur], n, hipMemcpyHostToDevice, s));
        QR_HIP_CHECK(hipEventRecord(pin.ev[cur], s));
        used[cur] = true;
        off += n;
        cur ^= 1;
    }
    QR_HIP_CHECK(hipStreamSynchronize(s));
    return QR_OK;
}

// ------------------------------------------------------------ raw format ---
const uint16_t kMark = uint16_t('M') | uint16_t('I') << 8;  // the bytes "MI", little endian

int raw_header(Source &src, bool &swab, uint64_t h[4], uint64_t &lp) {
    unsigned char m[2];
    if (src.read(m, 2) != 2) return fail(QR_ERR_FORMAT, "raw read: file too short");
    if (m[0] == 'M' && m[1] == 'I') swab = false;
    else if (m[0] == 'I' && m[1] == 'M') swab = true;
    else return fail(QR_ERR_FORMAT, "raw read: bad byte-order mark");
    uint64_t v[4];  // storage, rows, cols, indptr length
    if (src.read(v, sizeof v) != sizeof v) return fail(QR_ERR_FORMAT, "raw read: file too short");
    if (swab)
        for (auto &x : v) x = __builtin_bswap64(x);
    if (v[0] != 0) return fail(QR_ERR_FORMAT, "raw read: storage is not CSR");
    h[0] = v[1];  // rows
    h[1] = v[2];  // cols
    lp = v[3];    // indptr length
    if (lp != v[1] + 1) return fail(QR_ERR_FORMAT, "raw read: indptr length is not rows + 1");
    return QR_OK;
}

// --------------------------------------------------------- MatrixMarket ---
// Rust's Display for f64: the shortest digits that read back to the same
// value, written positionally (never with an exponent); "NaN", "inf", "-inf";
// -0.0 is "-0".  Writes at most 330 characters.
char *put_f64(char *p, double v) {
    if (std::isnan(v)) {
        memcpy(p, "NaN", 3);
        return p + 3;
    }
    if (std::isinf(v)) {
        if (v < 0) *p++ = '-';
        memcpy(p, "inf", 3);
        return p + 3;
    }
    char b[48];
    const auto r = std::to_chars(b, b + sizeof b, v, std::chars_format::scientific);
    const char *q = b;
    if (*q == '-') {
        *p++ = '-';
        ++q;
    }
    char dig[32];
    int nd = 0;
    while (*q != 'e') {
        if (*q != '.') dig[nd++] = *q;
        ++q;
    }
    ++q;
    const bool eneg = *q == '-';
    ++q;
    int ex = 0;
    while (q < r.ptr) ex = ex * 10 + (*q++ - '0');
    if (eneg) ex = -ex;
    // value = d1.d2...dn x 10^ex
    if (nd == 1 && dig[0] == '0') {
        *p++ = '0';
    } else if (ex >= nd - 1) {
        memcpy(p, dig, nd);
        p += nd;
        for (int k = 0; k < ex - (nd - 1); ++k) *p++ = '0';
    } else if (ex >= 0) {
        memcpy(p, dig, ex + 1);
        p += ex + 1;
        *p++ = '.';
        memcpy(p, dig + ex + 1, nd - ex - 1);
        p += nd - ex - 1;
    } else {
        *p++ = '0';
        *p++ = '.';
        for (int k = 0; k < -ex - 1; ++k) *p++ = '0';
        memcpy(p, dig, nd);
        p += nd;
    }
    return p;
}

char *put_u64(char *p, uint64_t v) { return std::to_chars(p, p + 24, v).ptr; }

const char *kSymNames[4] = {"general", "symmetric", "hermitian", "skew-symmetric"};

// entries kept by a symmetric mode: the lower triangle (row >= col), strictly
// (row > col) for skew-symmetric
__device__ __host__ inline bool mm_keep(int sym, uint64_t row, uint64_t col) {
    return sym == 0 || (sym == 3 ? row > col : row >= col);
}

__global__ __launch_bounds__(256) void k_count_kept(uint64_t nrows, uint64_t nnz, int sym,
                                                    const uint64_t *__restrict__ ip,
                                                    const uint64_t *__restrict__ ix,
                                                    unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[256];
    uint64_t n = 0;
    for (uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x; r < nrows;
         r += uint64_t(gridDim.x) * 256) {
        const uint64_t e1 = ip[r + 1] < nnz ? ip[r + 1] : nnz;
        for (uint64_t e = ip[r] < e1 ? ip[r] : e1; e < e1; ++e) n += mm_keep(sym, r, ix[e]);
    }
    part[threadIdx.x] = n;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

__global__ void k_zero_ull(unsigned long long *p) { *p = 0; }

// formats rows [r0, r1) whose entries sit in the host arrays at e - ebase
void format_rows(std::string &out, int sym, uint64_t r0, uint64_t r1, const uint64_t *ip,
                 uint64_t ebase, const uint64_t *ix, const double *d) {
    char line[800];
    for (uint64_t r = r0; r < r1; ++r) {
        for (uint64_t e = ip[r]; e < ip[r + 1]; ++e) {
            const uint64_t c = ix[e - ebase];
            if (!mm_keep(sym, r, c)) continue;
            char *p = put_u64(line, r + 1);
            *p++ = ' ';
            p = put_u64(p, c + 1);
            *p++ = ' ';
            p = put_f64(p, d[2 * (e - ebase)]);
            *p++ = ' ';
            p = put_f64(p, d[2 * (e - ebase) + 1]);
            *p++ = '\n';
            out.append(line, size_t(p - line));
        }
    }
}

// ---- reading
struct Triplets {
    std::vector<uint64_t> key;  // row << 32 | col, 0-based
    std::vector<double> val;    // re, im
};

inline const char *skip_ws(const char *p, const char *e) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    return p;
}

bool parse_u64(const char *&p, const char *e, uint64_t &v) {
    p = skip_ws(p, e);
    if (p < e && *p == '+') ++p;
    const auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
}

bool parse_f64(const char *&p, const char *e, double &v) {
    p = skip_ws(p, e);
    if (p < e && *p == '+') ++p;
    const auto r = std::from_chars(p, e, v, std::chars_format::general);
    if (r.ec != std::errc() && r.ec != std::errc::result_out_of_range) return false;
    p = r.ptr;
    return true;
}

std::string lower(std::string s) {
    for (auto &c : s) c = char(std::tolower((unsigned char)c));
    return s;
}

__global__ __launch_bounds__(256) void k_iota(uint64_t n, uint64_t *__restrict__ v) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ __launch_bounds__(256) void k_run_heads(uint64_t n, const uint64_t *__restrict__ key,
                                                   uint64_t *__restrict__ head) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i > n) return;
    head[i] = i < n && (i == 0 || key[i] != key[i - 1]);
}

// each run of equal keys summed in its (file) order into slot pos[i]
__global__ __launch_bounds__(256) void k_sum_runs(uint64_t n, const uint64_t *__restrict__ key,
                                                  const uint64_t *__restrict__ idx,
                                                  const double2 *__restrict__ val,
                                                  const uint64_t *__restrict__ pos,
                                                  uint64_t *__restrict__ ukey,
                                                  uint64_t *__restrict__ ix,
                                                  double2 *__restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n || (i > 0 && key[i] == key[i - 1])) return;
    double2 s = val[idx[i]];
    for (uint64_t j = i + 1; j < n && key[j] == key[i]; ++j) {
        const double2 v = val[idx[j]];
        s = make_double2(s.x + v.x, s.y + v.y);
    }
    const uint64_t p = pos[i];
    ukey[p] = key[i];
    ix[p] = key[i] & 0xFFFFFFFFull;
    out[p] = s;
}

// indptr[r] = first unique key with row >= r
__global__ __launch_bounds__(256) void k_row_starts(uint64_t nrows, uint64_t nu,
                                                    const uint64_t *__restrict__ ukey,
                                                    uint64_t *__restrict__ indptr) {
    const uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r > nrows) return;
    if (r == nrows) {  // (r << 32 would wrap for 2^32 rows)
        indptr[r] = nu;
        return;
    }
    const uint64_t k = r << 32;
    uint64_t lo = 0, hi = nu;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ukey[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    indptr[r] = lo;
}

unsigned blocks_of(uint64_t n) { return unsigned((n + 255) / 256); }

// RAII device buffer
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

}  // namespace

// ------------------------------------------------------------ raw entry ---
int io_write_raw(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz, const uint64_t *ip,
                 const uint64_t *ix, const double *d, hipStream_t s) {
    auto sink = open_sink(path);
    if (!sink) return fail(QR_ERR_IO, std::string("raw write: ") + path + ": " + os_reason());
    Pinned pin;
    if (!pin.ok) return fail(QR_ERR_NOMEM, "raw write: pinned staging buffers");
    bool ok = sink->write(reinterpret_cast<const char *>(&kMark), 2);
    const uint64_t h[4] = {0, rows, cols, rows + 1};
    ok = ok && sink->write(reinterpret_cast<const char *>(h), sizeof h);
    auto put = [&](const char *p, size_t n) { return sink->write(p, n) ? QR_OK : fail(QR_ERR_IO, "raw write: write failed"); };
    if (!ok) return fail(QR_ERR_IO, "raw write: write failed");
    int rc = stream_out(pin, ip, (rows + 1) * 8, Pinned::kCap, s, put);
    if (rc == QR_OK) rc = put(reinterpret_cast<const char *>(&nnz), 8);
    if (rc == QR_OK) rc = stream_out(pin, ix, nnz * 8, Pinned::kCap, s, put);
    if (rc == QR_OK) rc = put(reinterpret_cast<const char *>(&nnz), 8);
    if (rc == QR_OK) rc = stream_out(pin, d, nnz * 16, Pinned::kCap, s, put);
    if (rc != QR_OK) return rc;
    if (!sink->finish()) return fail(QR_ERR_IO, std::string("raw write: ") + path + ": " + os_reason());
    return QR_OK;
}

int io_read_raw_info(const char *path, uint64_t *rows, uint64_t *cols, uint64_t *nnz) {
    auto src = open_source(path);
    if (!src) return fail(QR_ERR_IO, std::string("raw read: ") + path + ": " + os_reason());
    bool swab;
    uint64_t h[4], lp;
    int rc = raw_header(*src, swab, h, lp);
    if (rc != QR_OK) return rc;
    // skip indptr to read the two array lengths
    std::vector<char> junk(1 << 20);
    for (uint64_t left = lp * 8; left;) {
        const size_t n = size_t(std::min<uint64_t>(left, junk.size()));
        if (src->read(junk.data(), n) != n) return fail(QR_ERR_FORMAT, "raw read: file too short");
        left -= n;
    }
    uint64_t li;
    if (src->read(&li, 8) != 8) return fail(QR_ERR_FORMAT, "raw read: file too short");
    if (swab) li = __builtin_bswap64(li);
    *rows = h[0];
    *cols = h[1];
    *nnz = li;
    return QR_OK;
}

int io_read_raw(const char *path, uint64_t rows, uint64_t nnz, uint64_t *ip, uint64_t *ix, double *d,
                hipStream_t s) {
    auto src = open_source(path);
    if (!src) return fail(QR_ERR_IO, std::string("raw read: ") + path + ": " + os_reason());
    bool swab;
    uint64_t h[4], lp;
    int rc = raw_header(*src, swab, h, lp);
    if (rc != QR_OK) return rc;
    if (h[0] != rows) return fail(QR_ERR_ARG, "raw read: rows differ from the file's");
    Pinned pin;
    if (!pin.ok) return fail(QR_ERR_NOMEM, "raw read: pinned staging buffers");
    rc = stream_in(pin, *src, ip, lp * 8, swab, s, "indptr");
    uint64_t len;
    if (rc == QR_OK) {
        if (src->read(&len, 8) != 8) return fail(QR_ERR_FORMAT, "raw read: file too short");
        if (swab) len = __builtin_bswap64(len);
        if (len != nnz) return fail(QR_ERR_ARG, "raw read: nnz differs from the file's");
        rc = stream_in(pin, *src, ix, len * 8, swab, s, "indices");
    }
    if (rc == QR_OK) {
        if (src->read(&len, 8) != 8) return fail(QR_ERR_FORMAT, "raw read: file too short");
        if (swab) len = __builtin_bswap64(len);
        if (len != nnz) return fail(QR_ERR_FORMAT, "raw read: data and indices lengths differ");
        rc = stream_in(pin, *src, d, len * 16, swab, s, "data");
    }
    return rc;
}

// --------------------------------------------------- MatrixMarket entry ---
int io_write_mm(const char *path, uint64_t rows, uint64_t cols, uint64_t nnz, const uint64_t *d_ip,
                const uint64_t *d_ix, const double *d_d, int sym, hipStream_t s) {
    if (sym < 0 || sym > 3) return fail(QR_ERR_ARG, "matrixmarket write: unknown symmetry mode");
    // host indptr (rows + 1 words), and the written entry count
    std::vector<uint64_t> ip(rows + 1);
    QR_HIP_CHECK(hipMemcpyAsync(ip.data(), d_ip, (rows + 1) * 8, hipMemcpyDeviceToHost, s));
    uint64_t kept = nnz;
    if (sym != 0) {
        DevBuf cnt;
        QR_HIP_CHECK(cnt.alloc(sizeof(unsigned long long)));
        hipLaunchKernelGGL(k_zero_ull, dim3(1), dim3(1), 0, s, cnt.as<unsigned long long>());
        const unsigned g = std::max(1u, std::min(blocks_of(rows), 4096u));
        hipLaunchKernelGGL(k_count_kept, dim3(g), dim3(256), 0, s, rows, nnz, sym, d_ip, d_ix,
                           cnt.as<unsigned long long>());
        QR_HIP_CHECK(hipGetLastError());
        unsigned long long k = 0;
        QR_HIP_CHECK(hipMemcpyAsync(&k, cnt.p, sizeof k, hipMemcpyDeviceToHost, s));
        QR_HIP_CHECK(hipStreamSynchronize(s));
        kept = k;
    }
    QR_HIP_CHECK(hipStreamSynchronize(s));
    if (ip[0] != 0 || ip[rows] != nnz) return fail(QR_ERR_ARG, "matrixmarket write: indptr does not span nnz");
    for (uint64_t r = 0; r < rows; ++r)
        if (ip[r + 1] < ip[r]) return fail(QR_ERR_ARG, "matrixmarket write: indptr decreases");

    auto sink = open_sink(path);
    if (!sink) return fail(QR_ERR_IO, std::string("matrixmarket write: ") + path + ": " + os_reason());
    std::string head = std::string("%%MatrixMarket matrix coordinate complex ") + kSymNames[sym] +
                       "\n% written by sprs\n";
    char nb[80];
    char *p = put_u64(nb, rows);
    *p++ = ' ';
    p = put_u64(p, cols);
    *p++ = ' ';
    p = put_u64(p, kept);
    *p++ = '\n';
    head.append(nb, size_t(p - nb));
    if (!sink->write(head.data(), head.size()))
        return fail(QR_ERR_IO, "matrixmarket write: write failed");

    // row chunks of at most kChunk entries (a longer row is a chunk of its own)
    constexpr uint64_t kChunk = Pinned::kCap / 16;
    Pinned pin_i, pin_d;
    if (!pin_i.ok || !pin_d.ok) return fail(QR_ERR_NOMEM, "matrixmarket write: pinned staging buffers");
    std::vector<std::pair<uint64_t, uint64_t>> chunks;  // row ranges
    for (uint64_t r = 0; r < rows;) {
        uint64_t r1 = r + 1;
        while (r1 < rows && ip[r1 + 1] - ip[r] <= kChunk) ++r1;
        chunks.emplace_back(r, r1);
        r = r1;
    }
    const unsigned T = io_threads();
    std::vector<std::string> parts(T);
    std::vector<uint64_t> big_ix;
    std::vector<double> big_d;
    int cur = 0;
    auto fetch = [&](size_t c, int slot) -> int {
        const uint64_t e0 = ip[chunks[c].first], n = ip[chunks[c].second] - e0;
        if (n > kChunk) return QR_OK;  // copied synchronously below
        QR_HIP_CHECK(hipMemcpyAsync(pin_i.buf[slot], d_ix + e0, n * 8, hipMemcpyDeviceToHost, s));
        QR_HIP_CHECK(hipMemcpyAsync(pin_d.buf[slot], d_d + 2 * e0, n * 16, hipMemcpyDeviceToHost, s));
        QR_HIP_CHECK(hipEventRecord(pin_i.ev[slot], s));
        return QR_OK;
    };
    if (!chunks.empty()) {
        int rc = fetch(0, 0);
        if (rc != QR_OK) return rc;
    }
    for (size_t c = 0; c < chunks.size(); ++c) {
        if (c + 1 < chunks.size()) {
            const int rc = fetch(c + 1, cur ^ 1);
            if (rc != QR_OK) return rc;
        }
        const uint64_t r0 = chunks[c].first, r1 = chunks[c].second;
        const uint64_t e0 = ip[r0], n = ip[r1] - e0;
        const uint64_t *hix;
        const double *hd;
        if (n > kChunk) {  // one long row
            QR_HIP_CHECK(hipStreamSynchronize(s));
            big_ix.resize(n);
            big_d.resize(2 * n);
            QR_HIP_CHECK(hipMemcpy(big_ix.data(), d_ix + e0, n * 8, hipMemcpyDeviceToHost));
            QR_HIP_CHECK(hipMemcpy(big_d.data(), d_d + 2 * e0, n * 16, hipMemcpyDeviceToHost));
            hix = big_ix.data();
            hd = big_d.data();
        } else {
            QR_HIP_CHECK(hipEventSynchronize(pin_i.ev[cur]));
            hix = reinterpret_cast<const uint64_t *>(pin_i.buf[cur]);
            hd = reinterpret_cast<const double *>(pin_d.buf[cur]);
        }
        // split the chunk's rows into T parts of about equal entries
        std::vector<uint64_t> cut(T + 1, r1);
        cut[0] = r0;
        for (unsigned t = 1; t < T; ++t) {
            const uint64_t target = e0 + n * t / T;
            cut[t] = uint64_t(std::lower_bound(ip.begin() + r0, ip.begin() + r1, target) - ip.begin());
            cut[t] = std::max(cut[t], cut[t - 1]);
        }
        parallel_for(T, [&](unsigned t) {
            parts[t].clear();
            format_rows(parts[t], sym, cut[t], std::max(cut[t], cut[t + 1]), ip.data(), e0, hix, hd);
        });
        for (unsigned t = 0; t < T; ++t)
            if (!sink->write(parts[t].data(), parts[t].size())) {
                (void)hipStreamSynchronize(s);
                return fail(QR_ERR_IO, "matrixmarket write: write failed");
            }
        cur ^= 1;
    }
    QR_HIP_CHECK(hipStreamSynchronize(s));
    if (!sink->finish())
        return fail(QR_ERR_IO, std::string("matrixmarket write: ") + path + ": " + os_reason());
    return QR_OK;
}

int io_read_mm(const char *path, qr_mm_reader **out, hipStream_t s) {
    auto src = open_source(path);
    if (!src) return fail(QR_ERR_IO, std::string("matrixmarket_read: ") + path + ": " + os_reason());
    std::string text;
    {
        std::vector<char> buf(size_t(8) << 20);
        for (;;) {
            const size_t n = src->read(buf.data(), buf.size());
            text.append(buf.data(), n);
            if (n < buf.size()) break;
        }
        if (src->bad) return fail(QR_ERR_IO, std::string("matrixmarket_read: ") + path + ": read error");
    }
    const char *p = text.data(), *end = p + text.size();
    auto next_line = [&](const char *&q) {
        const char *e = static_cast<const char *>(memchr(q, '\n', size_t(end - q)));
        const char *line_end = e ? e : end;
        std::string line(q, line_end);
        q = e ? e + 1 : end;
        return line;
    };
    // header: %%MatrixMarket matrix coordinate <field> <symmetry>
    std::string hl = lower(next_line(p));
    char f[5][64] = {};
    if (sscanf(hl.c_str(), "%63s %63s %63s %63s %63s", f[0], f[1], f[2], f[3], f[4]) != 5 ||
        strcmp(f[0], "%%matrixmarket") != 0 || strcmp(f[1], "matrix") != 0)
        return fail(QR_ERR_FORMAT, "matrixmarket_read: bad header line");
    if (strcmp(f[2], "coordinate") != 0)
        return fail(QR_ERR_FORMAT, "matrixmarket_read: only coordinate matrices are supported");
    const bool complex_field = strcmp(f[3], "complex") == 0;
    if (!complex_field && strcmp(f[3], "real") != 0 && strcmp(f[3], "integer") != 0)
        return fail(QR_ERR_FORMAT, std::string("matrixmarket_read: unsupported field ") + f[3]);
    int sym = -1;
    for (int k = 0; k < 4; ++k)
        if (strcmp(f[4], kSymNames[k]) == 0) sym = k;
    if (sym < 0) return fail(QR_ERR_FORMAT, std::string("matrixmarket_read: unknown symmetry ") + f[4]);
    // comments and blank lines, then the size line
    uint64_t rows = 0, cols = 0, nnz = 0;
    for (;;) {
        if (p >= end) return fail(QR_ERR_FORMAT, "matrixmarket_read: missing size line");
        std::string line = next_line(p);
        const char *a = skip_ws(line.data(), line.data() + line.size());
        if (a == line.data() + line.size() || *a == '%') continue;
        const char *le = line.data() + line.size();
        if (!parse_u64(a, le, rows) || !parse_u64(a, le, cols) || !parse_u64(a, le, nnz))
            return fail(QR_ERR_FORMAT, "matrixmarket_read: bad size line");
        break;
    }
    if (rows > (uint64_t(1) << 32) || cols > (uint64_t(1) << 32))
        return fail(QR_ERR_SIZE, "matrixmarket_read: more than 2^32 rows or columns");
    // entries, parsed by T threads over line-aligned pieces of the text
    const unsigned T = io_threads();
    std::vector<const char *> cut(T + 1, end);
    cut[0] = p;
    for (unsigned t = 1; t < T; ++t) {
        const char *c = p + size_t(end - p) * t / T;
        if (c < cut[t - 1]) c = cut[t - 1];
        const char *nl = c > p ? static_cast<const char *>(memchr(c - 1, '\n', size_t(end - c + 1))) : c;
        cut[t] = c == p ? p : (nl ? nl + 1 : end);
    }
    std::vector<Triplets> part(T);
    std::vector<std::string> err(T);
    parallel_for(T, [&](unsigned t) {
        const char *q = cut[t], *e = std::max(cut[t], cut[t + 1]);
        Triplets &tp = part[t];
        while (q < e) {
            const char *le = static_cast<const char *>(memchr(q, '\n', size_t(e - q)));
            if (!le) le = e;
            const char *a = skip_ws(q, le);
            if (a < le && *a != '%') {
                uint64_t i, j;
                double re = 0, im = 0;
                if (!parse_u64(a, le, i) || !parse_u64(a, le, j) || !parse_f64(a, le, re) ||
                    (complex_field && !parse_f64(a, le, im))) {
                    err[t] = "matrixmarket_read: bad entry line";
                    return;
                }
                if (i < 1 || i > rows || j < 1 || j > cols) {
                    err[t] = "matrixmarket_read: entry index out of range";
                    return;
                }
                const uint64_t r = i - 1, c = j - 1;
                tp.key.push_back(r << 32 | c);
                tp.val.push_back(re);
                tp.val.push_back(im);
                if (sym != 0 && r != c) {  // the mirrored entry: v, -v or conj v
                    tp.key.push_back(c << 32 | r);
                    tp.val.push_back(sym == 3 ? -re : re);
                    // a real or integer field has no imaginary part to negate
                    tp.val.push_back(!complex_field ? 0.0 : (sym == 1 ? im : -im));
                }
            }
            q = le < e ? le + 1 : e;
        }
    });
    uint64_t nexp = 0, nfile = 0;
    for (unsigned t = 0; t < T; ++t) {
        if (!err[t].empty()) return fail(QR_ERR_FORMAT, err[t]);
        nexp += part[t].key.size();
    }
    text.clear();
    text.shrink_to_fit();
    {
        // the file's entry lines: a diagonal line gave one entry, any other
        // line of a symmetric file two
        uint64_t diag = 0;
        for (unsigned t = 0; t < T; ++t)
            for (uint64_t k : part[t].key) diag += (k >> 32) == (k & 0xFFFFFFFFull);
        nfile = sym == 0 ? nexp : diag + (nexp - diag) / 2;
    }
    if (nfile != nnz)
        return fail(QR_ERR_FORMAT, "matrixmarket_read: entry count differs from the size line");

    // device: sort by key (stable), sum runs, row starts
    const uint64_t n = nexp;
    DevBuf k_in, k_out, i_in, i_out, vals, head, pos, ukey, tmp;
    QR_HIP_CHECK(k_in.alloc(n * 8));
    QR_HIP_CHECK(k_out.alloc(n * 8));
    QR_HIP_CHECK(i_in.alloc(n * 8));
    QR_HIP_CHECK(i_out.alloc(n * 8));
    QR_HIP_CHECK(vals.alloc(n * 16));
    {
        uint64_t off = 0;
        for (unsigned t = 0; t < T; ++t) {
            const uint64_t m = part[t].key.size();
            QR_HIP_CHECK(hipMemcpyAsync(k_in.as<uint64_t>() + off, part[t].key.data(), m * 8,
                                        hipMemcpyHostToDevice, s));
            QR_HIP_CHECK(hipMemcpyAsync(vals.as<double>() + 2 * off, part[t].val.data(), m * 16,
                                        hipMemcpyHostToDevice, s));
            off += m;
        }
        QR_HIP_CHECK(hipStreamSynchronize(s));
        part.clear();
    }
    std::unique_ptr<qr_mm_reader> rd(new qr_mm_reader);
    rd->rows = rows;
    rd->cols = cols;
    QR_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&rd->indptr), (rows + 1) * 8));
    uint64_t nu = 0;
    if (n > 0) {
        hipLaunchKernelGGL(k_iota, dim3(blocks_of(n)), dim3(256), 0, s, n, i_in.as<uint64_t>());
        size_t sort_bytes = 0, scan_bytes = 0;
        QR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k_in.as<uint64_t>(),
                                                        k_out.as<uint64_t>(), i_in.as<uint64_t>(),
                                                        i_out.as<uint64_t>(), n, 0, 64, s));
        QR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint64_t *)nullptr,
                                                      (uint64_t *)nullptr, n + 1, s));
        QR_HIP_CHECK(tmp.alloc(std::max(sort_bytes, scan_bytes)));
        QR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, sort_bytes, k_in.as<uint64_t>(),
                                                        k_out.as<uint64_t>(), i_in.as<uint64_t>(),
                                                        i_out.as<uint64_t>(), n, 0, 64, s));
        QR_HIP_CHECK(head.alloc((n + 1) * 8));
        QR_HIP_CHECK(pos.alloc((n + 1) * 8));
        hipLaunchKernelGGL(k_run_heads, dim3(blocks_of(n + 1)), dim3(256), 0, s, n,
                           k_out.as<uint64_t>(), head.as<uint64_t>());
        QR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan_bytes, head.as<uint64_t>(),
                                                      pos.as<uint64_t>(), n + 1, s));
        QR_HIP_CHECK(hipMemcpyAsync(&nu, pos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
        QR_HIP_CHECK(hipStreamSynchronize(s));
        QR_HIP_CHECK(ukey.alloc(nu * 8));
        QR_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&rd->indices), nu * 8));
        QR_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&rd->data), nu * 16));
        hipLaunchKernelGGL(k_sum_runs, dim3(blocks_of(n)), dim3(256), 0, s, n, k_out.as<uint64_t>(),
                           i_out.as<uint64_t>(), vals.as<double2>(), pos.as<uint64_t>(),
                           ukey.as<uint64_t>(), rd->indices, reinterpret_cast<double2 *>(rd->data));
        QR_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_row_starts, dim3(blocks_of(rows + 1)), dim3(256), 0, s, rows, nu,
                       ukey.as<uint64_t>(), rd->indptr);
    QR_HIP_CHECK(hipGetLastError());
    QR_HIP_CHECK(hipStreamSynchronize(s));
    rd->nnz = nu;
    *out = rd.release();
    return QR_OK;
}

}  // namespace qr
