// Plans streamed from column files: scans that read chunk windows from storage while earlier
// windows are scanned (the Q1 plan and star plans).
//
// The reference runs query1 as a per-chunk pipeline from SSD (usecases/ssb/queries/
// 01.hpp:350-424): each worker thread reads one block per column (VarLoader, core/access.hpp:
// 189-260), decompresses it and launches the probe kernel, skipping pruned blocks without
// reading them (skip_step, 01.hpp:368-378); the host blocks on every read and synchronizes on
// every chunk.
//
// Here the unit is a window of consecutive chunks (tens of MiB), so reads are large and a
// window's scan is a single fused launch (golap_q1_run, golap_star_run). Three stages overlap:
//   reader thread   pread window i+1 into pinned staging slot (i+1) % 2 (O_DIRECT, 4 KiB-aligned
//                   runs of unpruned chunks cut into 8 MiB pieces read by up to 8 threads at
//                   once; pruned chunks are never read);
//   copy stream     hipMemcpyAsync window i's runs into device slot i % 2, catalog and zone maps
//                   rebased to the window;
//   context stream  the plan's fused kernel over window i-1 (waits on its copy event).
// Device slot d is refilled only after the synchronous scan of the window it held has returned;
// staging slot s is refilled only after the copy that read it has completed (an event per slot).
#include <fcntl.h>
#include <unistd.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
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
#include "star_host.hpp"

using namespace golap_impl;
using namespace golap_impl::starhost;
using namespace golap_file;
using Lock = std::lock_guard<std::recursive_mutex>;

namespace {

constexpr int MAXC = 8;  // column files of one streamed plan
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
    uint64_t base[MAXC];      // 4 KiB-aligned start of the window's pages, relative to the data region
    uint64_t bytes[MAXC];     // end of the window's pages - base
    bool all_kept;            // every chunk of the window is scanned: the scan gets no chunk mask
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

// The column files of one plan, opened and their catalogs read; closed when it goes.
struct FileSet {
    ColumnFile f[MAXC];
    int n = 0;
    ~FileSet() {
        for (int k = 0; k < n; ++k) {
            if (f[k].fd >= 0) close(f[k].fd);
            if (f[k].bfd >= 0) close(f[k].bfd);
        }
    }
};

int open_files(golap_ctx *ctx, const char *const *paths, int n, FileSet &fs, const std::string &what) {
    for (int k = 0; k < n; ++k) {
        ColumnFile &c = fs.f[k];
        fs.n = k + 1;
        c.path = paths[k];
        std::string err;
        const int cfd = open(c.path.c_str(), O_RDONLY);
        if (cfd < 0) return fail(ctx, GOLAP_EINVAL, what + ": cannot open " + c.path);
        const bool ok = read_catalog(cfd, c.h, c.fh, err);
        close(cfd);
        if (!ok) return fail(ctx, GOLAP_ECORRUPT, err + " (" + c.path + ")");
        c.fd = open(c.path.c_str(), O_RDONLY | O_DIRECT);  // -1 on file systems without O_DIRECT
        c.direct = c.fd >= 0;
        c.bfd = open(c.path.c_str(), O_RDONLY);
        if (c.bfd < 0) return fail(ctx, GOLAP_EINVAL, what + ": cannot open " + c.path);
    }
    for (int k = 1; k < n; ++k) {
        const golap_host_column &a = fs.f[0].h, &b = fs.f[k].h;
        if (b.n_rows != a.n_rows || b.page_rows != a.page_rows || b.chunk_rows != a.chunk_rows)
            return fail(ctx, GOLAP_EINVAL, what + ": columns are not row-aligned");
    }
    return GOLAP_OK;
}

// Called per window with the window's device columns (one per file, catalogs rebased to the
// window), its chunk count and its device chunk mask (null when every chunk of the window is kept);
// runs the plan's scan synchronously.
using WindowScan = std::function<int(golap_column *const *cols, uint64_t n_chunks, const uint16_t *dmask)>;

// Streams the chunks with keep[ch] set through `scan` in windows of W chunks (0: about 64 MiB of
// pages per window); windows without kept chunks are counted in st.windows_skipped and their
// chunks in *skipped_chunks, nothing of them is read.
int stream_windows(golap_ctx *ctx, FileSet &fs, const std::vector<uint8_t> &keep, uint64_t W, golap_stream_stats &st,
                   uint64_t *skipped_chunks, const WindowScan &scan, const std::string &what) {
    const int nc = fs.n;
    ColumnFile *f = fs.f;
    const golap_host_column &H = f[0].h;
    const uint64_t n_chunks = H.n_chunks, ppc = H.chunk_rows / H.page_rows;
    if (!W) {
        uint64_t total = 0;
        for (int k = 0; k < nc; ++k) total += f[k].fh.comp_bytes;
        const uint64_t per_chunk = std::max<uint64_t>(1, total / std::max<uint64_t>(1, n_chunks));
        W = std::max<uint64_t>(1, (64ull << 20) / per_chunk);
    }
    W = std::max<uint64_t>(1, std::min<uint64_t>(W, std::max<uint64_t>(1, n_chunks)));
    st.window_chunks = (uint32_t)W;
    std::vector<Window> wins;
    uint64_t cap[MAXC] = {}, stage_cap[MAXC] = {};
    for (uint64_t c0 = 0; c0 < n_chunks; c0 += W) {
        Window w{};
        w.c0 = c0;
        w.c1 = std::min<uint64_t>(n_chunks, c0 + W);
        for (uint64_t ch = w.c0; ch < w.c1; ++ch) {
            if (!keep[ch]) continue;
            if (!w.runs.empty() && w.runs.back().c1 == ch) w.runs.back().c1 = ch + 1;
            else w.runs.push_back({ch, ch + 1});
        }
        w.all_kept = w.runs.size() == 1 && w.runs[0].c0 == w.c0 && w.runs[0].c1 == w.c1;
        if (w.runs.empty()) {
            st.windows_skipped++;
            *skipped_chunks += w.c1 - w.c0;
            continue;
        }
        for (int k = 0; k < nc; ++k) {
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
    golap_column *dcol[SLOTS][MAXC] = {};
    uint16_t *dmask[SLOTS] = {nullptr, nullptr};
    uint8_t *stage[SLOTS][MAXC] = {};
    hipEvent_t copied[SLOTS] = {}, staged[SLOTS] = {};
    hipStream_t cs = nullptr;
    std::vector<uint16_t> hmask[SLOTS];
    auto release = [&] {
        if (cs) (void)hipStreamSynchronize(cs);
        (void)hipStreamSynchronize(ctx->stream);
        for (int s = 0; s < SLOTS; ++s) {
            for (int k = 0; k < nc; ++k) {
                if (dcol[s][k]) golap_column_free(ctx, dcol[s][k]);
                if (stage[s][k]) (void)hipHostFree(stage[s][k]);
            }
            if (dmask[s]) dev_release(ctx, dmask[s]);
            if (copied[s]) (void)hipEventDestroy(copied[s]);
            if (staged[s]) (void)hipEventDestroy(staged[s]);
        }
        if (cs) (void)hipStreamDestroy(cs);
    };
    auto fail_release = [&](int code, const std::string &msg) {
        release();
        return fail(ctx, code, what + ": " + msg);
    };
    if (!wins.empty()) {
        const uint64_t nslots = std::min<uint64_t>(SLOTS, wins.size());
        for (uint64_t s = 0; s < nslots; ++s) {
            for (int k = 0; k < nc; ++k) {
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
            for (int k = 0; k < nc; ++k) {
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
    auto run_scan = [&](uint64_t i) -> int {
        const uint64_t s = i % SLOTS;
        GOLAP_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, copied[s], 0));
        if (int e = scan(dcol[s], wins[i].c1 - wins[i].c0, wins[i].all_kept ? nullptr : dmask[s])) return e;
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
        // catalogs first (each waits for its own small copies), then the page data, so the data
        // copies run behind the scan of window i - 1
        hipError_t e = hipSuccess;
        for (int k = 0; k < nc && status == GOLAP_OK; ++k) {
            golap_host_column wc;
            window_catalog(f[k].h, w.c0, w.c1, w.base[k], wc);
            status = set_column_window(ctx, dcol[s][k], wc, w.bytes[k], cs);
        }
        if (status) break;
        for (int k = 0; k < nc; ++k) {
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
        if (i >= 1) status = run_scan(i - 1);
    }
    if (status == GOLAP_OK && !wins.empty()) status = run_scan(wins.size() - 1);
    stop_reader();
    st.direct_io = 1;
    for (int k = 0; k < nc; ++k)
        if (!f[k].direct) st.direct_io = 0;
    release();
    if (status != GOLAP_OK) {
        if (!msg.empty()) return fail(ctx, status, what + ": " + msg);
        return status;
    }
    return GOLAP_OK;
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
    FileSet fs;
    if (int r = open_files(ctx, plan->paths, 4, fs, "q1_run_files")) return r;
    const ColumnFile *f = fs.f;
    const uint64_t n_chunks = f[0].h.n_chunks;
    if (plan->host_chunk_mask && plan->host_chunk_mask_len < n_chunks)
        return fail(ctx, GOLAP_EINVAL, "q1_run_files: host_chunk_mask shorter than the chunk count");
    // pruning: the host mask and the zone maps of the discount, quantity and price columns
    std::vector<uint8_t> keep(n_chunks, 1);
    for (uint64_t ch = 0; ch < n_chunks; ++ch) {
        if (plan->host_chunk_mask && !plan->host_chunk_mask[ch]) keep[ch] = 0;
        if (plan->zone_prune &&
            !(zone_hit(f[1].h, ch, plan->disc_lo, plan->disc_hi) && zone_hit(f[2].h, ch, plan->qty_lo, plan->qty_hi) &&
              zone_hit(f[3].h, ch, plan->price_lo, plan->price_hi)))
            keep[ch] = 0;
    }
    auto scan = [&](golap_column *const *cols, uint64_t nch, const uint16_t *dmask) -> int {
        golap_q1_plan qp{};
        qp.date = cols[0];
        qp.disc = cols[1];
        qp.qty = cols[2];
        qp.price = cols[3];
        qp.date_map = plan->date_map;
        qp.disc_lo = plan->disc_lo;
        qp.disc_hi = plan->disc_hi;
        qp.qty_lo = plan->qty_lo;
        qp.qty_hi = plan->qty_hi;
        qp.price_lo = plan->price_lo;
        qp.price_hi = plan->price_hi;
        qp.chunk_begin = 0;
        qp.chunk_end = nch;
        qp.dev_chunk_mask = dmask;
        golap_q1_result r{};
        if (int e = golap_q1_run(ctx, &qp, &r)) return e;
        res->sum += r.sum;
        res->matches += r.matches;
        res->chunks_run += r.chunks_run;
        res->chunks_pruned += r.chunks_pruned;
        res->comp_bytes += r.comp_bytes;
        res->kernel_ms += r.kernel_ms;
        return GOLAP_OK;
    };
    if (int r = stream_windows(ctx, fs, keep, plan->window_chunks, st, &res->chunks_pruned, scan, "q1_run_files"))
        return r;
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return GOLAP_OK;
}

int golap_star_run_files(golap_ctx *ctx, const golap_star_file_plan *plan, golap_star_result *res,
                         golap_stream_stats *stats, int32_t *group_keys, int64_t *sums, uint64_t *counts,
                         uint64_t max_groups) {
    if (!ctx || !plan || !res || !plan->value_path || plan->n_dims < 1 || plan->n_dims > GOLAP_STAR_MAX_DIMS)
        return GOLAP_EINVAL;
    const uint32_t nd = plan->n_dims;
    const char *paths[GOLAP_STAR_MAX_DIMS + 1];
    for (uint32_t d = 0; d < nd; ++d) {
        if (!plan->fk_paths[d] || !plan->maps[d]) return GOLAP_EINVAL;
        paths[d] = plan->fk_paths[d];
    }
    paths[nd] = plan->value_path;
    const auto t0 = std::chrono::steady_clock::now();
    memset(res, 0, sizeof(*res));
    golap_stream_stats st{};
    Lock l(ctx->mu);
    if (int r = set_device(ctx)) return r;
    // dense group domain: attributes in dimension order, first = most significant
    std::vector<int32_t> lo;
    std::vector<uint32_t> card;
    uint64_t dom = 1;
    for (uint32_t d = 0; d < nd; ++d)
        if (plan->group_card[d]) {
            lo.push_back(plan->group_lo[d]);
            card.push_back(plan->group_card[d]);
            dom *= plan->group_card[d];
            if (dom >= (1ull << 31)) return fail(ctx, GOLAP_EINVAL, "star_run_files: group domain too large");
        }
    std::vector<uint64_t> stride(card.size());
    for (int a = (int)card.size() - 1, sv = 1; a >= 0; --a) {
        stride[a] = (uint64_t)sv;
        sv *= (int)card[a];
    }
    FileSet fs;
    if (int r = open_files(ctx, paths, (int)nd + 1, fs, "star_run_files")) return r;
    const uint64_t n_chunks = fs.f[0].h.n_chunks;
    if (plan->host_chunk_mask && plan->host_chunk_mask_len < n_chunks)
        return fail(ctx, GOLAP_EINVAL, "star_run_files: host_chunk_mask shorter than the chunk count");
    std::vector<uint8_t> keep(n_chunks, 1);
    if (plan->host_chunk_mask)
        for (uint64_t ch = 0; ch < n_chunks; ++ch) keep[ch] = plan->host_chunk_mask[ch] != 0;
    // each window's groups added into the dense result
    std::vector<int64_t> hs(dom, 0);
    std::vector<uint64_t> hc(dom, 0);
    const size_t na = card.size();
    std::vector<int32_t> wk(dom * std::max<size_t>(1, na));
    std::vector<int64_t> ws(dom);
    std::vector<uint64_t> wcnt(dom);
    auto scan = [&](golap_column *const *cols, uint64_t nch, const uint16_t *dmask) -> int {
        golap_star_plan sp{};
        sp.n_dims = nd;
        for (uint32_t d = 0; d < nd; ++d)
            sp.dims[d] = golap_star_dim{cols[d], plan->maps[d], plan->group_lo[d], plan->group_card[d]};
        sp.value = cols[nd];
        sp.chunk_begin = 0;
        sp.chunk_end = nch;
        sp.dev_chunk_mask = dmask;
        sp.agg_mode = plan->agg_mode;
        golap_star_result r{};
        if (int e = golap_star_run(ctx, &sp, &r, wk.data(), ws.data(), wcnt.data(), dom)) return e;
        for (uint64_t g = 0; g < r.n_groups; ++g) {
            uint64_t idx = 0;
            for (size_t a = 0; a < na; ++a) idx += (uint64_t)(wk[g * na + a] - lo[a]) * stride[a];
            hs[idx] += ws[g];
            hc[idx] += wcnt[g];
        }
        res->chunks_run += r.chunks_run;
        res->chunks_pruned += r.chunks_pruned;
        res->comp_bytes += r.comp_bytes;
        res->kernel_ms += r.kernel_ms;
        res->driver_dim = r.driver_dim;
        res->lds_group_slots = r.lds_group_slots;
        return GOLAP_OK;
    };
    if (int r = stream_windows(ctx, fs, keep, plan->window_chunks, st, &res->chunks_pruned, scan, "star_run_files"))
        return r;
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return emit_groups_host(ctx, dom, hs.data(), hc.data(), lo, card, stride, &res->n_groups, &res->total_sum,
                            group_keys, sums, counts, max_groups, &res->rows_matched);
}

}  // extern "C"
