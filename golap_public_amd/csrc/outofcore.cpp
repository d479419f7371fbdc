// Q1 streamed from column files: scans that read chunk windows from storage while earlier
// windows are scanned.
//
// The reference runs query1 as a per-chunk pipeline from SSD (usecases/ssb/queries/
// 01.hpp:350-424): each worker thread reads one block per column (VarLoader, core/access.hpp:
// 189-260), decompresses it and launches the probe kernel, skipping pruned blocks without
// reading them (skip_step, 01.hpp:368-378); the host blocks on every read and synchronizes on
// every chunk.
//
// Here the unit is a window of consecutive chunks (tens of MiB), so reads are large and a
// window's scan is a single fused launch of golap_q1_run. Three stages overlap:
//   reader thread   pread window i+1 into pinned staging slot (i+1) % 2 (O_DIRECT, 4 KiB-aligned
//                   runs of unpruned chunks cut into 8 MiB pieces read by up to 8 threads at
//                   once; pruned chunks are never read);
//   copy stream     hipMemcpyAsync window i's runs into device slot i % 2, catalog and zone maps
//                   rebased to the window;
//   context stream  the fused Q1 kernel over window i-1 (waits on its copy event).
// Device slot d is refilled only after the synchronous scan of the window it held has returned;
// staging slot s is refilled only after the copy that read it has completed (an event per slot).
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

using namespace golap_impl;
using namespace golap_file;
using Lock = std::lock_guard<std::recursive_mutex>;

namespace {

constexpr int NCOL = 4;
constexpr int SLOTS = 2;

uint64_t align_down(uint64_t v) { return v / ALIGN * ALIGN; }

// Byte range [lo, hi) covering pages [p0, p1) of a catalog.
void page_span(const golap_host_column &h, uint64_t p0, uint64_t p1, uint64_t &lo, uint64_t &hi) {
    lo = UINT64_MAX;
    hi = 0;
    for (uint64_t p = p0; p < p1; ++p) {
        lo = std::min<uint64_t>(lo, h.page_off[p]);
        hi = std::max<uint64_t>(hi, h.page_off[p] + h.page_csize[p]);
    }
    if (lo > hi) lo = hi = 0;
}

struct Run {
    uint64_t c0, c1;  // chunks
};

struct Window {
    uint64_t c0, c1;          // chunks [c0, c1)
    std::vector<Run> runs;    // maximal ranges of unpruned chunks
    uint64_t base[NCOL];      // 4 KiB-aligned start of the window's pages, relative to the data region
    uint64_t bytes[NCOL];     // end of the window's pages - base
};

struct ColumnFile {
    int fd = -1;                     // O_DIRECT, or -1 where the file system refuses it
    int bfd = -1;                    // buffered
    std::atomic<bool> direct{false};  // O_DIRECT reads in use (cleared by the first refused one)
    std::string path;
    FileHeader fh{};
    golap_host_column h;  // catalog only
};

// Catalog of chunks [c0, c1) of h with page offsets rebased by base.
void window_catalog(const golap_host_column &h, uint64_t c0, uint64_t c1, uint64_t base, golap_host_column &w) {
    const uint64_t ppc = h.chunk_rows / h.page_rows;
    const uint64_t p0 = c0 * ppc, p1 = std::min<uint64_t>(h.n_pages, c1 * ppc);
    w.type = h.type;
    w.elem = h.elem;
    w.page_rows = h.page_rows;
    w.chunk_rows = h.chunk_rows;
    w.n_rows = std::min<uint64_t>(h.n_rows, c1 * h.chunk_rows) - c0 * h.chunk_rows;
    w.n_pages = p1 - p0;
    w.n_chunks = c1 - c0;
    w.page_off.resize(w.n_pages);
    w.page_csize.assign(h.page_csize.begin() + p0, h.page_csize.begin() + p1);
    w.page_codec.assign(h.page_codec.begin() + p0, h.page_codec.begin() + p1);
    w.chunk_min.assign(h.chunk_min.begin() + c0, h.chunk_min.begin() + c1);
    w.chunk_max.assign(h.chunk_max.begin() + c0, h.chunk_max.begin() + c1);
    w.payload_bytes = 0;
    for (uint64_t p = 0; p < w.n_pages; ++p) {
        w.page_off[p] = h.page_off[p0 + p] - base;
        w.payload_bytes += w.page_csize[p];
    }
}

// Whether chunk ch of a column can hold a value in [lo, hi] (zone map).
bool zone_hit(const golap_host_column &h, uint64_t ch, int64_t lo, int64_t hi) {
    return !(h.chunk_max[ch] < lo || h.chunk_min[ch] > hi);
}

}  // namespace

extern "C" {

int golap_q1_run_files(golap_ctx *ctx, const golap_q1_file_plan *plan, golap_q1_result *res,
                       golap_stream_stats *stats) {
    if (!ctx || !plan || !res) return GOLAP_EINVAL;
    for (const char *p : plan->paths)
        if (!p) return GOLAP_EINVAL;
    const auto t0 = std::chrono::steady_clock::now();
    memset(res, 0, sizeof(*res));
    golap_stream_stats st{};
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;

    // ---- open the files, read the catalogs
    ColumnFile f[NCOL];
    auto close_all = [&] {
        for (auto &c : f) {
            if (c.fd >= 0) close(c.fd);
            if (c.bfd >= 0) close(c.bfd);
        }
    };
    for (int k = 0; k < NCOL; ++k) {
        f[k].path = plan->paths[k];
        std::string err;
        const int cfd = open(f[k].path.c_str(), O_RDONLY);
        if (cfd < 0) {
            close_all();
            return fail(ctx, GOLAP_EINVAL, "q1_run_files: cannot open " + f[k].path);
        }
        const bool ok = read_catalog(cfd, f[k].h, f[k].fh, err);
        close(cfd);
        if (!ok) {
            close_all();
            return fail(ctx, GOLAP_ECORRUPT, err + " (" + f[k].path + ")");
        }
        f[k].fd = open(f[k].path.c_str(), O_RDONLY | O_DIRECT);  // -1 on file systems without O_DIRECT
        f[k].direct = f[k].fd >= 0;
        f[k].bfd = open(f[k].path.c_str(), O_RDONLY);
        if (f[k].bfd < 0) {
            close_all();
            return fail(ctx, GOLAP_EINVAL, "q1_run_files: cannot open " + f[k].path);
        }
    }
    const golap_host_column &H = f[0].h;
    for (int k = 1; k < NCOL; ++k)
        if (f[k].h.n_rows != H.n_rows || f[k].h.page_rows != H.page_rows || f[k].h.chunk_rows != H.chunk_rows) {
            close_all();
            return fail(ctx, GOLAP_EINVAL, "q1_run_files: columns are not row-aligned");
        }
    const uint64_t n_chunks = H.n_chunks, ppc = H.chunk_rows / H.page_rows;
    if (plan->host_chunk_mask && plan->host_chunk_mask_len < n_chunks) {
        close_all();
        return fail(ctx, GOLAP_EINVAL, "q1_run_files: host_chunk_mask shorter than the chunk count");
    }

    // ---- pruning and windows
    std::vector<uint8_t> keep(n_chunks, 1);
    for (uint64_t ch = 0; ch < n_chunks; ++ch) {
        if (plan->host_chunk_mask && !plan->host_chunk_mask[ch]) keep[ch] = 0;
        if (plan->zone_prune &&
            !(zone_hit(f[1].h, ch, plan->disc_lo, plan->disc_hi) && zone_hit(f[2].h, ch, plan->qty_lo, plan->qty_hi) &&
              zone_hit(f[3].h, ch, plan->price_lo, plan->price_hi)))
            keep[ch] = 0;
    }
    uint64_t W = plan->window_chunks;
    if (!W) {
        uint64_t total = 0;
        for (auto &c : f) total += c.fh.comp_bytes;
        const uint64_t per_chunk = std::max<uint64_t>(1, total / std::max<uint64_t>(1, n_chunks));
        W = std::max<uint64_t>(1, (64ull << 20) / per_chunk);
    }
    W = std::max<uint64_t>(1, std::min<uint64_t>(W, std::max<uint64_t>(1, n_chunks)));
    st.window_chunks = (uint32_t)W;
    std::vector<Window> wins;
    uint64_t cap[NCOL] = {0, 0, 0, 0}, stage_cap[NCOL] = {0, 0, 0, 0};
    for (uint64_t c0 = 0; c0 < n_chunks; c0 += W) {
        Window w{};
        w.c0 = c0;
        w.c1 = std::min<uint64_t>(n_chunks, c0 + W);
        for (uint64_t ch = w.c0; ch < w.c1; ++ch) {
            if (!keep[ch]) continue;
            if (!w.runs.empty() && w.runs.back().c1 == ch) w.runs.back().c1 = ch + 1;
            else w.runs.push_back({ch, ch + 1});
        }
        if (w.runs.empty()) {
            st.windows_skipped++;
            res->chunks_pruned += w.c1 - w.c0;
            continue;
        }
        for (int k = 0; k < NCOL; ++k) {
            uint64_t lo, hi;
            page_span(f[k].h, w.c0 * ppc, std::min<uint64_t>(f[k].h.n_pages, w.c1 * ppc), lo, hi);
            w.base[k] = align_down(lo);
            w.bytes[k] = hi - w.base[k];
            cap[k] = std::max(cap[k], w.bytes[k]);
            // staging holds the window's 4 KiB-aligned file blocks, relative to base
            stage_cap[k] = std::max(stage_cap[k], align_up(hi) - w.base[k]);
        }
        wins.push_back(std::move(w));
    }

    // ---- device slots (catalog capacity of a full window), staging, streams, events
    golap_column *dcol[SLOTS][NCOL] = {};
    uint16_t *dmask[SLOTS] = {nullptr, nullptr};
    uint8_t *stage[SLOTS][NCOL] = {};
    hipEvent_t copied[SLOTS] = {}, staged[SLOTS] = {};
    hipStream_t cs = nullptr;
    std::vector<uint16_t> hmask[SLOTS];
    auto release = [&] {
        if (cs) (void)hipStreamSynchronize(cs);
        (void)hipStreamSynchronize(ctx->stream);
        for (int s = 0; s < SLOTS; ++s) {
            for (int k = 0; k < NCOL; ++k) {
                if (dcol[s][k]) golap_column_free(ctx, dcol[s][k]);
                if (stage[s][k]) (void)hipHostFree(stage[s][k]);
            }
            if (dmask[s]) dev_release(ctx, dmask[s]);
            if (copied[s]) (void)hipEventDestroy(copied[s]);
            if (staged[s]) (void)hipEventDestroy(staged[s]);
        }
        if (cs) (void)hipStreamDestroy(cs);
        close_all();
    };
    auto fail_release = [&](int code, const std::string &msg) {
        release();
        return fail(ctx, code, "q1_run_files: " + msg);
    };
    if (!wins.empty()) {
        const uint64_t nslots = std::min<uint64_t>(SLOTS, wins.size());
        for (uint64_t s = 0; s < nslots; ++s) {
            for (int k = 0; k < NCOL; ++k) {
                golap_host_column capc;
                window_catalog(f[k].h, 0, W, 0, capc);  // the first window is a full one
                if (int r = create_device_column(ctx, &capc, cap[k], nullptr, &dcol[s][k])) {
                    release();
                    return r;
                }
                if (hipHostMalloc(reinterpret_cast<void **>(&stage[s][k]), std::max<uint64_t>(stage_cap[k], ALIGN),
                                  hipHostMallocDefault) != hipSuccess)
                    return fail_release(GOLAP_ENOMEM, "pinned staging");
            }
            void *m = nullptr;
            if (int r = dev_alloc(ctx, W * 2, &m)) {
                release();
                return r;
            }
            dmask[s] = static_cast<uint16_t *>(m);
            if (hipEventCreateWithFlags(&copied[s], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&staged[s], hipEventDisableTiming) != hipSuccess)
                return fail_release(GOLAP_EHIP, "events");
        }
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return fail_release(GOLAP_EHIP, "stream");
    }

    // ---- reader thread: window i into staging slot i % SLOTS once the copy of window i - SLOTS
    // (the slot's last reader) has completed
    std::mutex m;
    std::condition_variable cv;
    uint64_t filled = 0, enqueued = 0;
    bool stop = false, read_err = false;
    hipError_t reader_e = hipSuccess;
    std::string read_msg;
    std::thread reader([&] {
        for (uint64_t i = 0; i < wins.size(); ++i) {
            const uint64_t s = i % SLOTS;
            if (i >= SLOTS) {
                {
                    std::unique_lock<std::mutex> lk(m);
                    cv.wait(lk, [&] { return enqueued > i - SLOTS || stop; });
                    if (stop) return;
                }
                const hipError_t ev = hipEventSynchronize(staged[s]);
                if (ev != hipSuccess) {
                    std::lock_guard<std::mutex> lk(m);
                    reader_e = ev;
                    read_err = true;
                    cv.notify_all();
                    return;
                }
            }
            const Window &w = wins[i];
            // the window's unpruned runs as reads of whole 4 KiB blocks (the file is padded to a
            // block boundary), cut into pieces read by a few threads at once
            std::vector<FileRead> reads;
            for (int k = 0; k < NCOL; ++k) {
                ColumnFile &c = f[k];
                for (const Run &run : w.runs) {
                    uint64_t lo, hi;
                    page_span(c.h, run.c0 * ppc, std::min<uint64_t>(c.h.n_pages, run.c1 * ppc), lo, hi);
                    if (hi == lo) continue;
                    const uint64_t a = align_down(lo), b = std::min<uint64_t>(align_up(hi), c.fh.file_bytes - c.fh.data_offset);
                    split_read(reads, c.fd, c.bfd, &c.direct, c.fh.data_offset + a, b - a, stage[s][k] + (a - w.base[k]));
                }
            }
            const auto r0 = std::chrono::steady_clock::now();
            const bool ok = read_parallel(reads);
            uint64_t nread = 0;
            for (const FileRead &r : reads) nread += r.len;
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
            std::lock_guard<std::mutex> lk(m);
            st.read_ms += ms;
            st.bytes_read += nread;
            if (!ok) {
                read_err = true;
                read_msg = "short read";
            } else {
                filled = i + 1;
            }
            cv.notify_all();
            if (!ok) return;
        }
    });
    auto stop_reader = [&] {
        {
            std::lock_guard<std::mutex> lk(m);
            stop = true;
            cv.notify_all();
        }
        reader.join();
    };

    // ---- scan window i - 1 while window i is copied
    int status = GOLAP_OK;
    std::string msg;
    auto scan = [&](uint64_t i) -> int {
        const uint64_t s = i % SLOTS;
        const Window &w = wins[i];
        GOLAP_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, copied[s], 0));
        golap_q1_plan qp{};
        qp.date = dcol[s][0];
        qp.disc = dcol[s][1];
        qp.qty = dcol[s][2];
        qp.price = dcol[s][3];
        qp.date_map = plan->date_map;
        qp.disc_lo = plan->disc_lo;
        qp.disc_hi = plan->disc_hi;
        qp.qty_lo = plan->qty_lo;
        qp.qty_hi = plan->qty_hi;
        qp.price_lo = plan->price_lo;
        qp.price_hi = plan->price_hi;
        qp.chunk_begin = 0;
        qp.chunk_end = w.c1 - w.c0;
        qp.dev_chunk_mask = dmask[s];
        golap_q1_result r{};
        if (int e = golap_q1_run(ctx, &qp, &r)) return e;
        res->sum += r.sum;
        res->matches += r.matches;
        res->chunks_run += r.chunks_run;
        res->chunks_pruned += r.chunks_pruned;
        res->comp_bytes += r.comp_bytes;
        res->kernel_ms += r.kernel_ms;
        st.windows++;
        return GOLAP_OK;
    };
    for (uint64_t i = 0; i < wins.size() && status == GOLAP_OK; ++i) {
        const uint64_t s = i % SLOTS;
        const Window &w = wins[i];
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return filled > i || read_err; });
            if (read_err) {
                status = reader_e != hipSuccess ? GOLAP_EHIP : GOLAP_ECORRUPT;
                msg = reader_e != hipSuccess ? hipGetErrorString(reader_e) : read_msg;
                break;
            }
        }
        // device slot s held window i - 2, whose scan returned in the previous iteration; the
        // four catalogs first (each waits for its own small copies), then the page data, so the
        // data copies run behind the scan of window i - 1
        hipError_t e = hipSuccess;
        for (int k = 0; k < NCOL && status == GOLAP_OK; ++k) {
            golap_host_column wc;
            window_catalog(f[k].h, w.c0, w.c1, w.base[k], wc);
            status = set_column_window(ctx, dcol[s][k], wc, w.bytes[k], cs);
        }
        if (status) break;
        for (int k = 0; k < NCOL; ++k) {
            uint8_t *d = dcol[s][k]->d_comp;
            for (const Run &run : w.runs) {
                uint64_t lo, hi;
                page_span(f[k].h, run.c0 * ppc, std::min<uint64_t>(f[k].h.n_pages, run.c1 * ppc), lo, hi);
                if (hi == lo) continue;
                const uint64_t off = lo - w.base[k];
                if (e == hipSuccess) e = hipMemcpyAsync(d + off, stage[s][k] + off, hi - lo, hipMemcpyHostToDevice, cs);
                st.bytes_copied += hi - lo;
            }
            // dword pairs read past the window's last page land in zeros
            if (e == hipSuccess) e = hipMemsetAsync(d + w.bytes[k], 0, COMP_TAIL, cs);
        }
        hmask[s].assign(w.c1 - w.c0, 0);
        for (uint64_t ch = w.c0; ch < w.c1; ++ch) hmask[s][ch - w.c0] = keep[ch];
        if (e == hipSuccess) e = hipMemcpyAsync(dmask[s], hmask[s].data(), hmask[s].size() * 2, hipMemcpyHostToDevice, cs);
        if (e == hipSuccess) e = hipEventRecord(staged[s], cs);
        if (e == hipSuccess) e = hipEventRecord(copied[s], cs);
        {
            std::lock_guard<std::mutex> lk(m);
            enqueued = i + 1;
            cv.notify_all();
        }
        if (e != hipSuccess) {
            status = GOLAP_EHIP;
            msg = hipGetErrorString(e);
            break;
        }
        if (i >= 1) status = scan(i - 1);
    }
    if (status == GOLAP_OK && !wins.empty()) status = scan(wins.size() - 1);
    stop_reader();
    st.direct_io = 1;
    for (auto &c : f)
        if (!c.direct) st.direct_io = 0;
    release();
    if (status != GOLAP_OK) {
        if (!msg.empty()) return fail(ctx, status, "q1_run_files: " + msg);
        return status;
    }
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return GOLAP_OK;
}

}  // extern "C"
