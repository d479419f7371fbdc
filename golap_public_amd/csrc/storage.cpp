// Column files and the SSD -> HBM loader.
//
// The reference stores compressed chunks in 4 KiB-aligned blocks (DenseWriter,
// core/access.hpp:116-157; BlockInfo offsets, core/comp.cuh:23-78) and reads them back either
// with GPUDirect Storage (cuFile, core/storage.hpp:42-217) or through pinned host memory
// (SSD2CPU2GPU, core/apps.cuh:315-320; VarLoader, core/access.hpp:189-435): pread into a host
// buffer, then cudaMemcpyAsync to the device, one chunk per step per worker.
//
// Here a column file holds one paged column: a header, its catalog and zone maps, then the
// column's page data as one 4 KiB-aligned region (pages back to back, so each chunk's pages
// are contiguous). The loader reads that region in large segments with O_DIRECT preads (each
// segment in 8 MiB pieces by up to 8 threads) into a ring of pinned buffers on a reader thread
// while the calling thread copies filled buffers to
// HBM with hipMemcpyAsync: disk reads and PCIe copies overlap, and each buffer is reused once
// its copy has completed (an event per buffer). Files on file systems without O_DIRECT are
// read with buffered preads instead (reported in the statistics).
#include <fcntl.h>
#include <unistd.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "column_file.hpp"
#include "common.hpp"
#include "host_column.hpp"

namespace golap_file {

namespace {
constexpr char MAGIC[8] = {'G', 'O', 'L', 'A', 'P', 'C', 'O', 'L'};
constexpr uint32_t VERSION = 1;
}  // namespace

uint64_t align_up(uint64_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }

uint64_t catalog_bytes(uint64_t n_pages, uint64_t n_chunks) { return n_pages * (8 + 4 + 1) + n_chunks * 16; }

bool write_all(int fd, const void *p, uint64_t n, uint64_t off) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    while (n) {
        const ssize_t w = pwrite(fd, b, std::min<uint64_t>(n, 1ull << 30), (off_t)off);
        if (w <= 0) return false;
        b += w;
        n -= (uint64_t)w;
        off += (uint64_t)w;
    }
    return true;
}

bool read_all(int fd, void *p, uint64_t n, uint64_t off) {
    uint8_t *b = static_cast<uint8_t *>(p);
    while (n) {
        const ssize_t r = pread(fd, b, std::min<uint64_t>(n, 1ull << 30), (off_t)off);
        if (r <= 0) return false;
        b += r;
        n -= (uint64_t)r;
        off += (uint64_t)r;
    }
    return true;
}

// Header and catalog of a column file into h (page data left empty); false with err set.
bool read_catalog(int fd, golap_host_column &h, FileHeader &fh, std::string &err) {
    if (!read_all(fd, &fh, sizeof(fh), 0)) {
        err = "column file: short header";
        return false;
    }
    if (memcmp(fh.magic, MAGIC, 8) || fh.version != VERSION) {
        err = "column file: not a golap column file";
        return false;
    }
    const uint64_t cat = catalog_bytes(fh.n_pages, fh.n_chunks);
    if (!fh.page_rows || !fh.chunk_rows || fh.chunk_rows % fh.page_rows || fh.data_offset % ALIGN ||
        fh.data_offset < sizeof(fh) + cat || fh.file_bytes != align_up(fh.data_offset + fh.comp_bytes) ||
        fh.n_pages != (fh.n_rows + fh.page_rows - 1) / fh.page_rows ||
        fh.n_chunks != (fh.n_rows + fh.chunk_rows - 1) / fh.chunk_rows || golap_host::type_size(fh.type) != fh.elem) {
        err = "column file: inconsistent header";
        return false;
    }
    h.type = fh.type;
    h.elem = fh.elem;
    h.page_rows = fh.page_rows;
    h.chunk_rows = (uint32_t)fh.chunk_rows;
    h.n_rows = fh.n_rows;
    h.n_pages = fh.n_pages;
    h.n_chunks = fh.n_chunks;
    h.payload_bytes = fh.payload_bytes;
    h.page_off.resize(fh.n_pages);
    h.page_csize.resize(fh.n_pages);
    h.page_codec.resize(fh.n_pages);
    h.chunk_min.resize(fh.n_chunks);
    h.chunk_max.resize(fh.n_chunks);
    uint64_t off = sizeof(fh);
    bool ok = read_all(fd, h.page_off.data(), fh.n_pages * 8, off);
    off += fh.n_pages * 8;
    ok = ok && read_all(fd, h.page_csize.data(), fh.n_pages * 4, off);
    off += fh.n_pages * 4;
    ok = ok && read_all(fd, h.page_codec.data(), fh.n_pages, off);
    off += fh.n_pages;
    ok = ok && read_all(fd, h.chunk_min.data(), fh.n_chunks * 8, off);
    off += fh.n_chunks * 8;
    ok = ok && read_all(fd, h.chunk_max.data(), fh.n_chunks * 8, off);
    if (!ok) {
        err = "column file: short catalog";
        return false;
    }
    for (uint64_t p = 0; p < fh.n_pages; ++p)
        if (h.page_off[p] + h.page_csize[p] > fh.comp_bytes) {
            err = "column file: page outside the data";
            return false;
        }
    return true;
}

void split_read(std::vector<FileRead> &out, int dfd, int bfd, std::atomic<bool> *direct, uint64_t at, uint64_t len,
                uint8_t *dst) {
    for (uint64_t x = 0; x < len; x += READ_PIECE)
        out.push_back({dfd, bfd, direct, at + x, std::min<uint64_t>(READ_PIECE, len - x), dst + x});
}

bool read_parallel(const std::vector<FileRead> &reads) {
    std::atomic<size_t> next{0};
    std::atomic<bool> ok{true};
    auto work = [&] {
        for (size_t i = next++; i < reads.size() && ok; i = next++) {
            const FileRead &r = reads[i];
            bool done = false;
            if (*r.direct) {
                done = read_all(r.dfd, r.dst, r.len, r.at);
                if (!done) *r.direct = false;
            }
            if (!done) done = read_all(r.bfd, r.dst, r.len, r.at);
            if (!done) ok = false;
        }
    };
    const unsigned n = (unsigned)std::min<size_t>(READ_THREADS, reads.size());
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < n; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    return ok;
}

}  // namespace golap_file

using namespace golap_impl;
using namespace golap_file;
using Lock = std::lock_guard<std::recursive_mutex>;

extern "C" {

int golap_host_column_save(const golap_host_column *h, const char *path) {
    if (!h || !path) return GOLAP_EINVAL;
    FileHeader fh{};
    memcpy(fh.magic, MAGIC, 8);
    fh.version = VERSION;
    fh.type = h->type;
    fh.elem = h->elem;
    fh.page_rows = h->page_rows;
    fh.chunk_rows = h->chunk_rows;
    fh.n_rows = h->n_rows;
    fh.n_pages = h->n_pages;
    fh.n_chunks = h->n_chunks;
    fh.comp_bytes = h->comp.size();
    fh.payload_bytes = h->payload_bytes;
    fh.data_offset = align_up(sizeof(fh) + catalog_bytes(h->n_pages, h->n_chunks));
    fh.file_bytes = align_up(fh.data_offset + fh.comp_bytes);
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return GOLAP_EINVAL;
    uint64_t off = 0;
    bool ok = write_all(fd, &fh, sizeof(fh), off);
    off += sizeof(fh);
    ok = ok && write_all(fd, h->page_off.data(), h->n_pages * 8, off);
    off += h->n_pages * 8;
    ok = ok && write_all(fd, h->page_csize.data(), h->n_pages * 4, off);
    off += h->n_pages * 4;
    ok = ok && write_all(fd, h->page_codec.data(), h->n_pages, off);
    off += h->n_pages;
    ok = ok && write_all(fd, h->chunk_min.data(), h->n_chunks * 8, off);
    off += h->n_chunks * 8;
    ok = ok && write_all(fd, h->chunk_max.data(), h->n_chunks * 8, off);
    ok = ok && write_all(fd, h->comp.data(), fh.comp_bytes, fh.data_offset);
    ok = ok && ftruncate(fd, (off_t)fh.file_bytes) == 0;  // zero tail to the 4 KiB boundary
    ok = close(fd) == 0 && ok;
    return ok ? GOLAP_OK : GOLAP_EINVAL;
}

int golap_host_column_load(const char *path, golap_host_column **out) {
    if (!path || !out) return GOLAP_EINVAL;
    *out = nullptr;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return GOLAP_EINVAL;
    std::unique_ptr<golap_host_column> h(new golap_host_column());
    FileHeader fh{};
    std::string err;
    bool ok = read_catalog(fd, *h, fh, err);
    if (ok) {
        h->comp.resize(fh.comp_bytes);
        ok = read_all(fd, h->comp.data(), fh.comp_bytes, fh.data_offset);
    }
    close(fd);
    if (!ok) return GOLAP_ECORRUPT;
    *out = h.release();
    return GOLAP_OK;
}

int golap_column_load_file(golap_ctx *ctx, const char *path, uint64_t segment_bytes, uint32_t n_buffers,
                           golap_column **out, golap_load_stats *stats) {
    if (!ctx || !path || !out) return GOLAP_EINVAL;
    *out = nullptr;
    if (stats) memset(stats, 0, sizeof(*stats));
    segment_bytes = segment_bytes ? align_up(segment_bytes) : (64ull << 20);
    n_buffers = std::min<uint32_t>(std::max<uint32_t>(n_buffers ? n_buffers : 4, 2), 16);
    const auto t0 = std::chrono::steady_clock::now();
    struct Fds {
        int dfd, bfd;  // O_DIRECT (-1 on file systems without it) and buffered
        ~Fds() {
            if (dfd >= 0) close(dfd);
            if (bfd >= 0) close(bfd);
        }
    } fds{open(path, O_RDONLY | O_DIRECT), open(path, O_RDONLY)};
    std::atomic<bool> direct{fds.dfd >= 0};
    if (fds.bfd < 0) return fail(ctx, GOLAP_EINVAL, std::string("load_file: cannot open ") + path);
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    // header and catalog through a pinned, aligned buffer (O_DIRECT needs aligned transfers)
    golap_host_column h;
    FileHeader fh{};
    {
        std::string err;
        const int cfd = open(path, O_RDONLY);
        if (cfd < 0) return fail(ctx, GOLAP_EINVAL, "load_file: cannot reopen");
        const bool ok = read_catalog(cfd, h, fh, err);
        close(cfd);
        if (!ok) return fail(ctx, GOLAP_ECORRUPT, err);
    }
    golap_column *col = nullptr;
    if (int r = create_device_column(ctx, &h, fh.comp_bytes, nullptr, &col)) return r;
    auto drop = [&](int code, const std::string &msg) {
        golap_column_free(ctx, col);
        return fail(ctx, code, msg);
    };
    // ring of pinned buffers: the reader fills buffer i % n while the copy of buffer i - n
    // has completed
    std::vector<uint8_t *> buf(n_buffers, nullptr);
    std::vector<hipEvent_t> done(n_buffers, nullptr);
    for (uint32_t i = 0; i < n_buffers; ++i) {
        if (hipHostMalloc(reinterpret_cast<void **>(&buf[i]), segment_bytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) {
            for (uint32_t j = 0; j <= i; ++j) {
                if (buf[j]) (void)hipHostFree(buf[j]);
                if (done[j]) (void)hipEventDestroy(done[j]);
            }
            return drop(GOLAP_ENOMEM, "load_file: pinned buffers");
        }
    }
    const uint64_t data_end = fh.data_offset + fh.comp_bytes;
    const uint64_t n_seg = (fh.comp_bytes + segment_bytes - 1) / segment_bytes;
    std::mutex m;
    std::condition_variable cv;
    uint64_t filled = 0, enqueued = 0;  // segments read; segments whose copy (and event) is enqueued
    bool stop = false, read_err = false;
    hipError_t reader_e = hipSuccess;
    double read_ms = 0;
    std::thread reader([&] {
        for (uint64_t t = 0; t < n_seg; ++t) {
            if (t >= n_buffers) {
                // buffer t % n is free once the copy of segment t - n (its last user) is done
                {
                    std::unique_lock<std::mutex> lk(m);
                    cv.wait(lk, [&] { return enqueued > t - n_buffers || stop; });
                    if (stop) return;
                }
                const hipError_t ev = hipEventSynchronize(done[t % n_buffers]);
                if (ev != hipSuccess) {
                    std::lock_guard<std::mutex> lk(m);
                    reader_e = ev;
                    read_err = true;
                    cv.notify_all();
                    return;
                }
            }
            const uint64_t off = fh.data_offset + t * segment_bytes;
            const uint64_t want = std::min<uint64_t>(segment_bytes, data_end - off);
            const auto r0 = std::chrono::steady_clock::now();
            // whole 4 KiB blocks (the file is padded to a block boundary), in pieces read by a
            // few threads at once
            std::vector<FileRead> reads;
            split_read(reads, fds.dfd, fds.bfd, &direct, off, align_up(want), buf[t % n_buffers]);
            const bool ok = read_parallel(reads);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
            std::lock_guard<std::mutex> lk(m);
            read_ms += ms;
            if (!ok) read_err = true;
            else filled = t + 1;
            cv.notify_all();
            if (!ok) return;
        }
    });
    hipError_t e = hipSuccess;
    for (uint64_t s = 0; s < n_seg && e == hipSuccess; ++s) {
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return filled > s || read_err; });
            if (read_err) break;
        }
        const uint64_t off = s * segment_bytes, n = std::min<uint64_t>(segment_bytes, fh.comp_bytes - off);
        e = hipMemcpyAsync(col->d_comp + off, buf[s % n_buffers], n, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(done[s % n_buffers], ctx->stream);
        std::lock_guard<std::mutex> lk(m);
        enqueued = s + 1;
        cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(m);
        stop = true;  // a reader still waiting for a buffer ends
        cv.notify_all();
    }
    reader.join();
    if (e == hipSuccess) e = reader_e;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    for (uint32_t i = 0; i < n_buffers; ++i) {
        (void)hipHostFree(buf[i]);
        (void)hipEventDestroy(done[i]);
    }
    if (e != hipSuccess) return drop(GOLAP_EHIP, std::string("load_file: ") + hipGetErrorString(e));
    if (read_err) return drop(GOLAP_ECORRUPT, "load_file: short read");
    if (stats) {
        stats->bytes = fh.comp_bytes;
        stats->segments = n_seg;
        stats->direct_io = direct;
        stats->read_ms = read_ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    *out = col;
    return GOLAP_OK;
}

}  // extern "C"
