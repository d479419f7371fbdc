// golap.hpp - C++ facade over the C ABI (golap_hip.h) that keeps the reference's names for the
// compressed-chunk -> filter -> join-probe -> aggregate path of nboeschen/golap_public, so plan
// code written against core/apps.cuh, core/comp.cuh, core/dev_structs.cuh, core/join.cuh,
// core/metadata.cuh and core/core.hpp reads the same. Header-only; link libgolap_hip.so.
//
// What differs from the reference, by design:
//  * errors throw golap::Error (the reference aborts in checkCudaErrors);
//  * compressed chunks stay resident in HBM, or stream from column files through pinned host
//    memory (query1_files); no cuFile;
//  * the fused plans (query1, star) take whole chunk ranges in one launch instead of one
//    kernel per chunk driven from host worker threads; the per-chunk Op protocol remains for
//    callers that step chunks themselves (ChunkDecompressor, Collector, TableLoader, load),
//    with the reference's streams and events (CStream, CEvent, Op::step(stream, parent_event)).
#pragma once
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../golap_hip.h"

namespace golap {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

namespace detail {
inline void check(int status, golap_ctx *ctx, const char *what) {
    if (status != GOLAP_OK)
        throw Error(status, std::string(what) + ": " + golap_status_string(status) + " (" +
                                (ctx ? golap_last_error(ctx) : "") + ")");
}
template <typename T>
constexpr uint32_t type_of() {
    if (std::is_same<T, int8_t>::value) return GOLAP_I8;
    if (std::is_same<T, uint8_t>::value) return GOLAP_U8;
    if (std::is_same<T, int16_t>::value) return GOLAP_I16;
    if (std::is_same<T, uint16_t>::value) return GOLAP_U16;
    if (std::is_same<T, int32_t>::value) return GOLAP_I32;
    if (std::is_same<T, uint32_t>::value) return GOLAP_U32;
    return GOLAP_I64;
}
}  // namespace detail

// The device context (cudaSetDevice + stream set-up of the reference, 01.hpp:61).
class Device {
public:
    explicit Device(int device = 0) { detail::check(golap_ctx_create(device, &ctx_), nullptr, "ctx_create"); }
    ~Device() {
        if (ctx_) golap_ctx_destroy(ctx_);
    }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    golap_ctx *get() const { return ctx_; }
    void sync() { detail::check(golap_sync(ctx_), ctx_, "sync"); }
    // golap::DEVICE_ALLOCATED (core/mem.hpp:14-15), per context here
    uint64_t allocated() const { return golap_device_allocated(ctx_); }

private:
    golap_ctx *ctx_ = nullptr;
};

template <typename T>
struct Tag {};

// DeviceMem (core/mem.hpp:136-219): owning device buffer.
class DeviceMem {
public:
    DeviceMem() = default;
    template <typename T>
    DeviceMem(Device &dev, Tag<T>, uint64_t num) : dev_(&dev), bytes_(num * sizeof(T)) {
        detail::check(golap_dev_alloc(dev.get(), bytes_, &ptr_), dev.get(), "dev_alloc");
    }
    DeviceMem(DeviceMem &&o) noexcept { *this = std::move(o); }
    DeviceMem &operator=(DeviceMem &&o) noexcept {
        std::swap(dev_, o.dev_);
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
        return *this;
    }
    ~DeviceMem() {
        if (ptr_ && dev_) golap_dev_free(dev_->get(), ptr_);
    }
    template <typename T>
    T *ptr() const { return static_cast<T *>(ptr_); }
    uint64_t size_bytes() const { return bytes_; }
    template <typename T>
    void copy_from(const T *host, uint64_t num) {
        detail::check(golap_h2d(dev_->get(), ptr_, host, num * sizeof(T)), dev_->get(), "h2d");
    }
    template <typename T>
    void copy_to(T *host, uint64_t num) const {
        detail::check(golap_d2h(dev_->get(), host, ptr_, num * sizeof(T)), dev_->get(), "d2h");
    }

private:
    Device *dev_ = nullptr;
    void *ptr_ = nullptr;
    uint64_t bytes_ = 0;
};

// BlockInfo (core/comp.cuh:23-32): one chunk of a compressed column.
struct BlockInfo {
    uint64_t offset;   // byte offset of the chunk's first page in the column buffer
    uint64_t size;     // bytes of the chunk's pages including alignment gaps
    uint64_t payload;  // bytes of the chunk's pages
    uint64_t tuples;   // rows in the chunk
};

// MinMaxMeta<T> (core/metadata.cuh:110-221): per-chunk zone map, host copy.
template <typename T>
struct MinMaxMeta {
    std::vector<int64_t> mins, maxs;
    // check_pred_host (metadata.cuh:195-200): chunk qualifies iff lo <= max && hi >= min
    bool check_pred_host(uint64_t chunk, T lo, T hi) const {
        return (int64_t)lo <= maxs[chunk] && (int64_t)hi >= mins[chunk];
    }
};

// CompInfo (core/comp.cuh:35-78): catalog of one HBM-resident compressed column.
struct CompInfo {
    uint64_t chunk_bytes = 0;   // uncompressed bytes per chunk
    uint64_t uncomp_bytes = 0;
    std::string comp_algo = "LZ4";
    uint64_t page_rows = 4096;  // rows per independently coded page (nvcomp's nvchunk role)
    uint32_t lz4_min_match = 4;
    uint32_t lz4_unit = 1;      // LZ4 sequence unit (1: byte-granular; 0: the element width, typed)
    uint32_t value_size = 0;    // bytes per value, set by prepare_compressed_device
    std::vector<BlockInfo> blocks;
    std::shared_ptr<golap_column> col;  // device column

    uint64_t get_comp_bytes() const {
        uint64_t s = 0;
        for (auto &b : blocks) s += b.payload;
        return s;
    }
    uint64_t max_uncomp_chunk() const { return chunk_bytes; }
};

inline uint32_t codec_of(const std::string &algo) {
    if (algo == "LZ4") return GOLAP_CODEC_LZ4;
    if (algo == "DELTA" || algo == "Cascaded") return GOLAP_CODEC_DBP;
    if (algo == "RLE") return GOLAP_CODEC_RLE;
    if (algo == "UNCOMPRESSED") return GOLAP_CODEC_RAW;
    if (algo == "AUTO") return GOLAP_CODEC_AUTO;
    throw Error(GOLAP_EINVAL, "unknown comp_algo " + algo);
}

// EqHistogram (core/metadata.cuh:225-382) and BloomMeta (:386-537): per-chunk metadata built
// on the device by prepare_compressed_device and checked by MetaChecker::add_hist / add_bloom.
template <typename T>
class EqHistogram {
public:
    explicit EqHistogram(uint64_t bucket_num) : bucket_num(bucket_num) {}
    uint64_t bucket_num;
    std::shared_ptr<golap_chunkmeta> meta;
};
template <typename T>
class BloomMeta {
public:
    BloomMeta(double p, uint64_t m) : p(p), m(m) {}
    double p;
    uint64_t m;  // filter bytes per chunk
    std::shared_ptr<golap_chunkmeta> meta;
};

// prepare_compressed_device (core/apps.cuh:257-275): encode `col` into row-aligned chunks of
// ci.chunk_bytes, upload to HBM, fill ci.blocks and the zone map (and, when given, the
// histogram and Bloom metadata of each chunk). Returns the encode time (ms).
namespace detail {
template <typename X>
struct same {  // keeps a parameter out of template argument deduction (so nullptr binds)
    using type = X;
};
}  // namespace detail
namespace detail {
// A device column c of num_tuples values into ci (catalog -> BlockInfo list), with the zone maps
// into minmax and the histogram / Bloom metadata built from the values at dev_values.
template <typename T>
void adopt_column(Device &dev, golap_column *c, uint64_t num_tuples, CompInfo &ci, MinMaxMeta<T> *minmax,
                  EqHistogram<T> *histogram, BloomMeta<T> *bloommeta, const void *dev_values) {
    golap_ctx *ctx = dev.get();
    ci.col = std::shared_ptr<golap_column>(c, [ctx](golap_column *p) { golap_column_free(ctx, p); });
    ci.uncomp_bytes = num_tuples * sizeof(T);
    ci.value_size = sizeof(T);
    golap_column_desc d;
    golap_column_info(c, &d);
    // catalog and zone maps back to the host
    std::vector<uint64_t> off(d.n_pages);
    std::vector<uint32_t> csz(d.n_pages);
    golap_d2h(ctx, off.data(), d.page_off, d.n_pages * 8);
    golap_d2h(ctx, csz.data(), d.page_csize, d.n_pages * 4);
    const uint64_t ppc = d.chunk_rows / d.page_rows;
    ci.blocks.clear();
    for (uint64_t ch = 0; ch < d.n_chunks; ++ch) {
        const uint64_t p0 = ch * ppc, p1 = std::min<uint64_t>(d.n_pages, p0 + ppc);
        BlockInfo b{off[p0], off[p1 - 1] + csz[p1 - 1] - off[p0], 0,
                    std::min<uint64_t>(d.chunk_rows, num_tuples - ch * d.chunk_rows)};
        for (uint64_t p = p0; p < p1; ++p) b.payload += csz[p];
        ci.blocks.push_back(b);
    }
    if (minmax) {
        minmax->mins.resize(d.n_chunks);
        minmax->maxs.resize(d.n_chunks);
        golap_d2h(ctx, minmax->mins.data(), d.chunk_min, d.n_chunks * 8);
        golap_d2h(ctx, minmax->maxs.data(), d.chunk_max, d.n_chunks * 8);
    }
    auto keep = [ctx](golap_chunkmeta *m) {
        return std::shared_ptr<golap_chunkmeta>(m, [ctx](golap_chunkmeta *q) { golap_chunkmeta_free(ctx, q); });
    };
    golap_chunkmeta *m = nullptr;
    if (histogram) {
        check(golap_eqhist_build(ctx, dev_values, type_of<T>(), num_tuples, d.chunk_rows,
                                 (uint32_t)histogram->bucket_num, &m),
              ctx, "eqhist_build");
        histogram->meta = keep(m);
    }
    if (bloommeta) {
        check(golap_bloommeta_build(ctx, dev_values, type_of<T>(), num_tuples, d.chunk_rows, bloommeta->p,
                                    bloommeta->m, &m),
              ctx, "bloommeta_build");
        bloommeta->meta = keep(m);
    }
}

template <typename T>
golap_encode_opts encode_opts_of(const CompInfo &ci) {
    golap_encode_opts o;
    golap_encode_opts_default(&o);
    o.codec = codec_of(ci.comp_algo);
    o.page_rows = (uint32_t)ci.page_rows;
    o.chunk_rows = (uint32_t)(ci.chunk_bytes / sizeof(T));
    o.lz4_min_match = ci.lz4_min_match;
    o.lz4_unit = ci.lz4_unit;
    return o;
}
}  // namespace detail

// prepare_compressed_device (core/apps.cuh:257-275) for host values: coded on the host, uploaded;
// returns the coding time in ms.
template <typename T>
double prepare_compressed_device(Device &dev, const T *col, uint64_t num_tuples, CompInfo &ci,
                                 typename detail::same<MinMaxMeta<T>>::type *minmax = nullptr,
                                 typename detail::same<EqHistogram<T>>::type *histogram = nullptr,
                                 typename detail::same<BloomMeta<T>>::type *bloommeta = nullptr) {
    const golap_encode_opts o = detail::encode_opts_of<T>(ci);
    golap_column *c = nullptr;
    double ms = 0;
    detail::check(golap_prepare_compressed_device(dev.get(), col, num_tuples, detail::type_of<T>(), &o, &c, &ms),
                  dev.get(), "prepare_compressed_device");
    golap_ctx *ctx = dev.get();
    if (histogram || bloommeta) {
        void *v = nullptr;
        detail::check(golap_dev_alloc(ctx, std::max<uint64_t>(num_tuples, 1) * sizeof(T), &v), ctx, "dev_alloc");
        std::shared_ptr<void> guard(v, [ctx](void *q) { golap_dev_free(ctx, q); });
        detail::check(golap_h2d(ctx, v, col, num_tuples * sizeof(T)), ctx, "h2d");
        detail::adopt_column<T>(dev, c, num_tuples, ci, minmax, histogram, bloommeta, v);
    } else {
        detail::adopt_column<T>(dev, c, num_tuples, ci, minmax, nullptr, nullptr, nullptr);
    }
    return ms;
}

// The same for values already in HBM (the reference's Column in device memory, coded by nvcomp
// there): coded on the device, byte-identical to the host path; returns the device time in ms.
// Call with the value type: prepare_compressed_device<int32_t>(dev, values, n, ci, &minmax).
template <typename T>
double prepare_compressed_device(Device &dev, const DeviceMem &dev_col, uint64_t num_tuples, CompInfo &ci,
                                 typename detail::same<MinMaxMeta<T>>::type *minmax = nullptr,
                                 typename detail::same<EqHistogram<T>>::type *histogram = nullptr,
                                 typename detail::same<BloomMeta<T>>::type *bloommeta = nullptr) {
    if (dev_col.size_bytes() < num_tuples * sizeof(T))
        throw Error(GOLAP_EINVAL, "prepare_compressed_device: buffer smaller than num_tuples values");
    const golap_encode_opts o = detail::encode_opts_of<T>(ci);
    golap_column *c = nullptr;
    double ms = 0;
    detail::check(golap_encode_device(dev.get(), dev_col.ptr<T>(), num_tuples, detail::type_of<T>(), &o, &c, &ms),
                  dev.get(), "encode_device");
    detail::adopt_column<T>(dev, c, num_tuples, ci, minmax, histogram, bloommeta, dev_col.ptr<T>());
    return ms;
}

// Groups of a group-by or star plan in key order: keys[i] holds one value per group attribute.
struct GroupResult {
    std::vector<std::vector<int32_t>> keys;
    std::vector<int64_t> sums;
    std::vector<uint64_t> counts;
};

// ---------------------------------------------------------------- streams and events (core/core.hpp:192-235)
// CStream: a stream of its own (golap_stream_create); .stream is what Op::step takes. A null
// golap_stream* anywhere means the context's own stream, where the fused plans run.
struct CStream {
    explicit CStream(Device &dev) : dev_(&dev) {
        detail::check(golap_stream_create(dev.get(), &stream), dev.get(), "stream_create");
    }
    CStream(const CStream &) = delete;
    CStream &operator=(const CStream &) = delete;
    ~CStream() {
        if (stream) golap_stream_destroy(stream);
    }
    // waits for the stream; throws on a corrupt page or key out of range raised by its work
    void sync() { detail::check(golap_stream_sync(stream), dev_->get(), "stream_sync"); }
    golap_stream *stream = nullptr;

private:
    Device *dev_;
};
struct CEvent {
    explicit CEvent(Device &dev) : dev_(&dev) {
        detail::check(golap_event_create(dev.get(), &event), dev.get(), "event_create");
    }
    CEvent(const CEvent &) = delete;
    CEvent &operator=(const CEvent &) = delete;
    ~CEvent() {
        if (event) golap_event_destroy(event);
    }
    void record(golap_stream *s) { detail::check(golap_event_record(event, s), dev_->get(), "event_record"); }
    void sync() { detail::check(golap_event_sync(event), dev_->get(), "event_sync"); }
    golap_event *event = nullptr;

private:
    Device *dev_;
};

// ---------------------------------------------------------------- Op protocol (core/core.hpp:94-124)
// step(stream, parent_event): enqueue the next chunk's work on stream once parent_event (may be
// null) has fired, then record this op's event on stream; false when exhausted. skip_step
// advances without work (a pruned chunk, 01.hpp:368-378) and records the event too, so waiters
// never see a stale one. The argument-less forms run one step on the context's stream and wait.
class Op {
public:
    explicit Op(Device &dev) : dev_(dev), ev_(dev), event(ev_.event) {}
    Op(const Op &) = delete;
    Op &operator=(const Op &) = delete;
    virtual ~Op() = default;
    virtual bool step(golap_stream *stream, golap_event *parent_event) = 0;
    virtual bool skip_step(golap_stream *stream, golap_event *parent_event) = 0;
    virtual void finish_step(golap_stream *stream) {
        detail::check(stream ? golap_stream_sync(stream) : golap_sync(dev_.get()), dev_.get(), "finish_step");
    }
    bool step() {
        const bool r = step(nullptr, nullptr);
        finish_step(nullptr);
        return r;
    }
    bool skip_step() { return skip_step(nullptr, nullptr); }

protected:
    void wait_for(golap_stream *stream, golap_event *parent_event) {
        if (parent_event)
            detail::check(golap_stream_wait_event(dev_.get(), stream, parent_event), dev_.get(), "stream_wait_event");
    }
    void done(golap_stream *stream) { ev_.record(stream); }
    Device &dev_;

private:
    CEvent ev_;

public:
    golap_event *event;          // recorded at the end of every step
    uint64_t last_produced = 0;  // bytes produced by the last step
};

// ChunkDecompressor (core/comp.cuh:151-216): decodes the next chunk of its block list into its
// own device buffer (decomp_buf) on the stream it is stepped on.
class ChunkDecompressor : public Op {
public:
    ChunkDecompressor(Device &dev, const CompInfo &ci, std::vector<uint64_t> block_idxs)
        : Op(dev), ci_(ci), idxs_(std::move(block_idxs)), decomp_buf(dev, Tag<uint8_t>{}, ci.chunk_bytes) {}
    using Op::skip_step;
    using Op::step;
    bool step(golap_stream *stream, golap_event *parent_event) override {
        if (cur_ >= idxs_.size()) return false;
        const uint64_t b = idxs_[cur_++];
        wait_for(stream, parent_event);
        detail::check(golap_decompress_chunks_async(dev_.get(), ci_.col.get(), b, b + 1, decomp_buf.ptr<void>(), stream),
                      dev_.get(), "decompress");
        last_produced = ci_.blocks[b].tuples * ci_.value_size;
        done(stream);
        return true;
    }
    bool skip_step(golap_stream *stream, golap_event *parent_event) override {
        if (cur_ >= idxs_.size()) return false;
        ++cur_;
        last_produced = 0;
        wait_for(stream, parent_event);
        done(stream);
        return true;
    }
    uint64_t current_block() const { return idxs_[cur_ - 1]; }
    const CompInfo &ci_;
    std::vector<uint64_t> idxs_;
    uint64_t cur_ = 0;
    DeviceMem decomp_buf;
};

// Collector (core/core.hpp:252-307): steps every child for the same chunk, each on its own
// stream behind the parent event; the collector's stream then waits for every child's event.
class Collector : public Op {
public:
    explicit Collector(Device &dev) : Op(dev) {}
    using Op::skip_step;
    using Op::step;
    // opstream null: the child runs on the stream the collector is stepped on
    void add_child(Op *op, golap_stream *opstream) { children.emplace_back(op, opstream); }
    void add(Op *child) { add_child(child, nullptr); }
    bool step(golap_stream *stream, golap_event *parent_event) override {
        bool res = false;
        for (auto &c : children) {
            golap_stream *cs = c.second ? c.second : stream;
            wait_for(cs, parent_event);
            res = c.first->step(cs, parent_event) || res;
            if (cs != stream) wait_for(stream, c.first->event);
        }
        done(stream);
        return res;
    }
    bool skip_step(golap_stream *stream, golap_event *parent_event) override {
        bool res = false;
        for (auto &c : children) {
            golap_stream *cs = c.second ? c.second : stream;
            res = c.first->skip_step(cs, parent_event) || res;
            if (cs != stream) wait_for(stream, c.first->event);
        }
        done(stream);
        return res;
    }
    std::vector<std::pair<Op *, golap_stream *>> children;
};

// DecompressEnv (core/apps.cuh:341-396) for one column of one worker's block slice: its blocks,
// its stream and event, and the decompressor that is its root op.
struct DecompressEnv {
    DecompressEnv(Device &dev, const std::vector<uint64_t> &all_block_idxs, uint64_t start, uint64_t past_end,
                  const CompInfo &ci)
        : myblock_idxs(all_block_idxs.begin() + start, all_block_idxs.begin() + past_end),
          gstream(dev), gevent(dev), root(dev, ci, myblock_idxs) {
        for (uint64_t b : myblock_idxs) myblocks.push_back(ci.blocks[b]);
    }
    std::vector<uint64_t> myblock_idxs;
    std::vector<BlockInfo> myblocks;
    CStream gstream;
    CEvent gevent;
    ChunkDecompressor root;
    DeviceMem &decomp_buf() { return root.decomp_buf; }
};

// TableLoader<BLOCKENV> (core/apps.cuh:611-630): one env per column, each column stepped on its
// env's stream by one Collector (rootop); rootstream/rootevent are the caller's.
template <typename BLOCKENV>
struct TableLoader {
    explicit TableLoader(Device &dev, uint32_t reserve = 4) : dev_(dev), rootop(dev), rootstream(dev), rootevent(dev) {
        blockenvs.reserve(reserve);
    }
    void add(const std::string &name, const std::vector<uint64_t> &all_block_idxs, uint64_t start,
             uint64_t past_end, const CompInfo &ci) {
        names.push_back(name);
        blockenvs.emplace_back(new BLOCKENV(dev_, all_block_idxs, start, past_end, ci));
        rootop.add_child(&blockenvs.back()->root, blockenvs.back()->gstream.stream);
    }
    BLOCKENV &at(const std::string &name) {
        for (size_t i = 0; i < names.size(); ++i)
            if (names[i] == name) return *blockenvs[i];
        throw Error(GOLAP_EINVAL, "TableLoader: no column " + name);
    }
    Device &dev_;
    std::vector<std::string> names;
    std::vector<std::unique_ptr<BLOCKENV>> blockenvs;
    Collector rootop;
    CStream rootstream;
    CEvent rootevent;
};

// load<BLOCKENV, T> (core/apps.cuh:682-793): worker_num threads, each with its env over a
// contiguous slice of the column's chunks, step their decompressor on their own stream and copy
// every decoded chunk to host_ptr + chunk * chunk_bytes (the SSD2GPU2CPU dataflow; chunks stay
// in HBM here). Chunks whose prune_check flag is 0 are skipped (skip_step) and not written.
// Returns the wall time in ms; pruned_chunks (may be null) gets the skipped count.
template <typename BLOCKENV, typename T>
double load(Device &dev, uint64_t worker_num, T *host_ptr, const CompInfo &ci, uint64_t &device_allocated,
            const uint16_t *prune_check = nullptr, uint64_t *pruned_chunks = nullptr) {
    const uint64_t nb = ci.blocks.size();
    worker_num = std::max<uint64_t>(1, std::min<uint64_t>(worker_num, std::max<uint64_t>(nb, 1)));
    std::vector<uint64_t> all(nb);
    for (uint64_t i = 0; i < nb; ++i) all[i] = i;
    std::vector<std::unique_ptr<BLOCKENV>> envs;
    for (uint64_t w = 0; w < worker_num; ++w)  // util::SliceSeq: contiguous, sizes differ by at most one
        envs.emplace_back(new BLOCKENV(dev, all, nb * w / worker_num, nb * (w + 1) / worker_num, ci));
    std::vector<std::thread> threads;
    std::vector<std::string> errors(worker_num);
    std::atomic<uint64_t> pruned{0};
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t w = 0; w < worker_num; ++w) {
        threads.emplace_back([&, w] {
            try {
                BLOCKENV &env = *envs[w];
                for (size_t round = 0; round < env.myblocks.size(); ++round) {
                    const uint64_t b = env.myblock_idxs[round];
                    if (prune_check && prune_check[b] == 0) {
                        env.root.skip_step(env.gstream.stream, env.gevent.event);
                        pruned.fetch_add(1, std::memory_order_relaxed);
                        continue;
                    }
                    env.root.step(env.gstream.stream, env.gevent.event);
                    detail::check(golap_stream_copy(dev.get(), env.gstream.stream,
                                                    reinterpret_cast<uint8_t *>(host_ptr) + b * ci.chunk_bytes,
                                                    env.decomp_buf().template ptr<void>(),
                                                    env.myblocks[round].tuples * ci.value_size),
                                  dev.get(), "load copy");
                    env.gevent.record(env.gstream.stream);
                }
                env.gstream.sync();
            } catch (const std::exception &e) {
                errors[w] = e.what();
            }
        });
    }
    for (auto &t : threads) t.join();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (auto &e : errors)
        if (!e.empty()) throw Error(GOLAP_ECORRUPT, "load: " + e);
    device_allocated = dev.allocated();
    if (pruned_chunks) *pruned_chunks = pruned.load();
    return ms;
}

// HashAggregate (core/dev_structs.cuh:157-253): a group-by accumulator that stays on the device
// across add() calls (each one HashAggregate::add per row, with an atomic SUM and a count).
// The reference takes the number of groups and its own group/result arrays; here the group
// domain is given by each key's lo and card (golap_group_key.dev_col is not used).
class HashAggregate {
public:
    HashAggregate(Device &dev, const std::vector<golap_group_key> &keys, uint32_t agg_mode = GOLAP_AGG_AUTO)
        : dev_(dev), n_keys_((uint32_t)keys.size()) {
        detail::check(golap_hashagg_create(dev.get(), keys.data(), n_keys_, agg_mode, &h_), dev.get(), "hashagg");
    }
    HashAggregate(const HashAggregate &) = delete;
    HashAggregate &operator=(const HashAggregate &) = delete;
    ~HashAggregate() {
        if (h_) golap_hashagg_destroy(h_);
    }
    // n rows of the key columns (device, one per key, the types given at construction) and a
    // value column, on stream (null: the context's stream)
    template <typename V>
    void add(const std::vector<const void *> &dev_keys, const V *dev_val, uint64_t n, golap_stream *stream = nullptr) {
        detail::check(golap_hashagg_add(h_, dev_keys.data(), dev_val, detail::type_of<V>(), n, stream), dev_.get(),
                      "hashagg_add");
    }
    // the groups with at least one row, in key order (waits for the context's stream: make it
    // wait for the adds on other streams first, e.g. through a Collector's event)
    GroupResult results(uint64_t max_groups) {
        std::vector<int32_t> k(max_groups * n_keys_);
        GroupResult g;
        g.sums.resize(max_groups);
        g.counts.resize(max_groups);
        uint64_t n = 0;
        detail::check(golap_hashagg_results(h_, &n, k.data(), g.sums.data(), g.counts.data(), max_groups), dev_.get(),
                      "hashagg_results");
        g.sums.resize(n);
        g.counts.resize(n);
        for (uint64_t i = 0; i < n; ++i) g.keys.emplace_back(k.begin() + i * n_keys_, k.begin() + (i + 1) * n_keys_);
        return g;
    }
    void clear() { detail::check(golap_hashagg_clear(h_), dev_.get(), "hashagg_clear"); }

private:
    Device &dev_;
    uint32_t n_keys_;
    golap_hashagg *h_ = nullptr;
};

// ---------------------------------------------------------------- joins (dev_structs.cuh, join.cuh)
// PredInfo<T, ConstPred> (dev_structs.cuh:21-33) as an inclusive range over a device column.
template <typename T>
struct PredInfo {
    const T *col;
    int64_t lo, hi;
};
template <typename T>
PredInfo<T> ConstPred(const T *col, T c) { return PredInfo<T>{col, (int64_t)c, (int64_t)c}; }

// HashMap(ht_size, table) (dev_structs.cuh:47-113) over a device key column.
class HashMap {
public:
    template <typename KEY>
    HashMap(Device &dev, uint64_t ht_size, const KEY *dev_table) : dev_(dev), ht_size(ht_size) {
        detail::check(golap_hashmap_create(dev.get(), ht_size, dev_table, detail::type_of<KEY>(), &m_), dev.get(),
                      "hashmap_create");
    }
    ~HashMap() {
        if (m_) golap_hashmap_free(dev_.get(), m_);
    }
    HashMap(const HashMap &) = delete;
    HashMap &operator=(const HashMap &) = delete;
    uint64_t filled() const {
        uint64_t f = 0;
        golap_hashmap_filled(dev_.get(), m_, &f);
        return f;
    }
    // probe n keys: rid (or payload) per key, UINT64_MAX when absent (HashMap::probe)
    template <typename KEY>
    void probe(const KEY *dev_keys, uint64_t n, uint64_t *dev_out) {
        detail::check(golap_hashmap_probe(dev_.get(), m_, dev_keys, detail::type_of<KEY>(), n, dev_out), dev_.get(),
                      "probe");
    }
    golap_hashmap *get() const { return m_; }
    Device &dev_;
    uint64_t ht_size;

private:
    golap_hashmap *m_ = nullptr;
};

namespace detail {
inline void collect_preds(std::vector<golap_pred> &) {}
template <typename T, typename... REST>
void collect_preds(std::vector<golap_pred> &v, const PredInfo<T> &p, const REST &...rest) {
    v.push_back(golap_pred{p.col, type_of<T>(), p.lo, p.hi});
    collect_preds(v, rest...);
}
}  // namespace detail

// hash_map_build_pred_other (core/join.cuh:131-135): insert every row passing all predicates.
template <typename... PREDS>
void hash_map_build_pred_other(HashMap &m, uint64_t num_tuples, const PREDS &...preds) {
    std::vector<golap_pred> v;
    detail::collect_preds(v, preds...);
    detail::check(golap_hashmap_build(m.dev_.get(), m.get(), num_tuples, v.data(), (uint32_t)v.size(), nullptr),
                  m.dev_.get(), "hash_map_build_pred_other");
}
// Same, storing a per-row int32 payload (a group attribute) beside each key.
template <typename... PREDS>
void hash_map_build_pred_payload(HashMap &m, uint64_t num_tuples, const int32_t *dev_payload, const PREDS &...preds) {
    std::vector<golap_pred> v;
    detail::collect_preds(v, preds...);
    detail::check(golap_hashmap_build(m.dev_.get(), m.get(), num_tuples, v.data(), (uint32_t)v.size(), dev_payload),
                  m.dev_.get(), "hash_map_build");
}
// hash_join_count (core/join.cuh:45-56)
template <typename KEY>
uint64_t hash_join_count(HashMap &m, const KEY *dev_probe, uint64_t num) {
    uint64_t hits = 0;
    detail::check(golap_hash_join_count(m.dev_.get(), m.get(), dev_probe, detail::type_of<KEY>(), num, &hits),
                  m.dev_.get(), "hash_join_count");
    return hits;
}

// HostHashMap (core/host_structs.hpp:9-78) over a host int32 key column; build inserts the rows
// passing every predicate (PredInfo over host columns), storing a payload (or the rid).
class HostHashMap {
public:
    HostHashMap(uint64_t ht_size, const int32_t *table) : ht_size(ht_size) {
        detail::check(golap_host_hashmap_create(ht_size, table, &m_), nullptr, "host_hashmap_create");
    }
    ~HostHashMap() {
        if (m_) golap_host_hashmap_free(m_);
    }
    HostHashMap(const HostHashMap &) = delete;
    HostHashMap &operator=(const HostHashMap &) = delete;
    template <typename... PREDS>
    void build(uint64_t num_tuples, const int32_t *payload, const PREDS &...preds) {
        std::vector<golap_pred> v;
        detail::collect_preds(v, preds...);
        std::vector<golap_host_pred> hp;
        for (const golap_pred &p : v) hp.push_back(golap_host_pred{p.dev_col, p.type, p.lo, p.hi});
        detail::check(golap_host_hashmap_build(m_, num_tuples, hp.data(), (uint32_t)hp.size(), payload, 0), nullptr,
                      "host_hashmap_build");
    }
    // HostHashMap::probe: payload (or rid) of the key, UINT64_MAX when absent
    uint64_t probe(int32_t key) const {
        uint64_t out = 0;
        detail::check(golap_host_hashmap_probe(m_, &key, 1, &out, nullptr), nullptr, "host_hashmap_probe");
        return out;
    }
    uint64_t filled() const {
        uint64_t f = 0;
        golap_host_hashmap_filled(m_, &f);
        return f;
    }
    golap_host_hashmap *get() const { return m_; }
    uint64_t ht_size;

private:
    golap_host_hashmap *m_ = nullptr;
};

// ---------------------------------------------------------------- pruning (metadata.cuh)
// MetaChecker (core/metadata.cuh:638-698): AND of zone-map predicates into one u16 chunk mask.
class MetaChecker {
public:
    MetaChecker(Device &dev, uint64_t chunk_num)
        : dev_(dev), chunk_num(chunk_num), combined(dev, Tag<uint16_t>{}, chunk_num),
          tmp(dev, Tag<uint16_t>{}, chunk_num) {
        detail::check(golap_memset(dev.get(), combined.ptr<void>(), 0, chunk_num * 2), dev.get(), "memset");
    }
    // check_mmmeta + combine_and (metadata.cuh:539-547, 632-636)
    template <typename T>
    void add_minmax(const CompInfo &ci, T lo, T hi) {
        uint16_t *dst = first_ ? combined.ptr<uint16_t>() : tmp.ptr<uint16_t>();
        detail::check(golap_minmax_check(dev_.get(), ci.col.get(), (int64_t)lo, (int64_t)hi, dst), dev_.get(),
                      "minmax_check");
        combine(dst);
    }
    // check_hist + combine_and (metadata.cuh:573-585, MetaChecker::add_hist :674-687)
    template <typename T>
    void add_hist(const EqHistogram<T> &h, T lo, T hi) {
        uint16_t *dst = first_ ? combined.ptr<uint16_t>() : tmp.ptr<uint16_t>();
        detail::check(golap_eqhist_check(dev_.get(), h.meta.get(), (int64_t)lo, (int64_t)hi, dst), dev_.get(),
                      "eqhist_check");
        combine(dst);
    }
    // check_bloom + combine_and (metadata.cuh:593-600, MetaChecker::add_bloom :689-702)
    template <typename T>
    void add_bloom(const BloomMeta<T> &b, T val) {
        uint16_t *dst = first_ ? combined.ptr<uint16_t>() : tmp.ptr<uint16_t>();
        detail::check(golap_bloommeta_check(dev_.get(), b.meta.get(), (int64_t)val, dst), dev_.get(),
                      "bloommeta_check");
        combine(dst);
    }
    // Join-key pruning (usecases/ssb/queries/01.hpp:246-270): collect_hashmap_keys of the
    // built map, check_mmmeta_multiple_pts of the foreign-key column against them, combine_and.
    void add_join_keys(const HashMap &map, const CompInfo &fk) {
        const uint64_t cap = std::max<uint64_t>(map.filled(), 1);
        DeviceMem keys(dev_, Tag<int32_t>{}, cap);
        uint64_t n = 0;
        detail::check(golap_hashmap_keys(dev_.get(), map.get(), keys.ptr<int32_t>(), cap, &n), dev_.get(),
                      "hashmap_keys");
        uint16_t *dst = first_ ? combined.ptr<uint16_t>() : tmp.ptr<uint16_t>();
        detail::check(golap_minmax_check_points(dev_.get(), fk.col.get(), keys.ptr<int32_t>(), n, dst), dev_.get(),
                      "minmax_check_points");
        combine(dst);
    }
    const uint16_t *mask() const { return first_ ? nullptr : combined.ptr<uint16_t>(); }
    // ANDs the flags just written into tmp into the combined mask (the first check wrote it).
    void combine(uint16_t *dst) {
        if (!first_)
            detail::check(golap_combine_and(dev_.get(), combined.ptr<uint16_t>(), dst, chunk_num), dev_.get(),
                          "combine_and");
        first_ = false;
    }
    std::vector<uint16_t> to_host() const {
        std::vector<uint16_t> h(chunk_num, 1);
        if (!first_) combined.copy_to(h.data(), chunk_num);
        return h;
    }
    Device &dev_;
    uint64_t chunk_num;

private:
    DeviceMem combined, tmp;
    bool first_ = true;
};

// ---------------------------------------------------------------- plans
// query1 (usecases/ssb/queries/01.hpp:53-449) fused: SUM(price*discount) over the chunk range.
struct Q1Result {
    int64_t sum;
    uint64_t matches, chunks_run, chunks_pruned, comp_bytes;
    double kernel_ms;
};
inline Q1Result query1(Device &dev, const CompInfo &orderdate, const CompInfo &discount, const CompInfo &quantity,
                       const CompInfo &extendedprice, const HashMap &date_map, int64_t disc_lo, int64_t disc_hi,
                       int64_t qty_lo, int64_t qty_hi, const MetaChecker *checker = nullptr,
                       int64_t price_lo = INT32_MIN, int64_t price_hi = INT32_MAX) {
    golap_q1_plan p{};
    p.date = orderdate.col.get();
    p.disc = discount.col.get();
    p.qty = quantity.col.get();
    p.price = extendedprice.col.get();
    p.date_map = date_map.get();
    p.disc_lo = disc_lo;
    p.disc_hi = disc_hi;
    p.qty_lo = qty_lo;
    p.qty_hi = qty_hi;
    p.price_lo = price_lo;
    p.price_hi = price_hi;
    p.chunk_begin = 0;
    p.chunk_end = orderdate.blocks.size();
    p.dev_chunk_mask = checker ? checker->mask() : nullptr;
    golap_q1_result r{};
    detail::check(golap_q1_run(dev.get(), &p, &r), dev.get(), "query1");
    return Q1Result{r.sum, r.matches, r.chunks_run, r.chunks_pruned, r.comp_bytes, r.kernel_ms};
}

// HashAggregate result of a grouping plan (dev_structs.cuh:157-253), canonical order.
// query2 / query3 (02.hpp, 03.hpp) fused star join + group-by.
struct StarDim {
    const CompInfo *fk;
    const HashMap *map;
    int32_t group_lo = 0;
    uint32_t group_card = 0;
};
inline GroupResult star_query(Device &dev, const std::vector<StarDim> &dims, const CompInfo &value,
                              const MetaChecker *checker = nullptr, uint64_t max_groups = 1 << 20) {
    golap_star_plan p{};
    p.n_dims = (uint32_t)dims.size();
    uint32_t na = 0;
    for (size_t i = 0; i < dims.size(); ++i) {
        p.dims[i] = golap_star_dim{dims[i].fk->col.get(), dims[i].map->get(), dims[i].group_lo, dims[i].group_card};
        na += dims[i].group_card ? 1 : 0;
    }
    p.value = value.col.get();
    p.chunk_end = value.blocks.size();
    p.dev_chunk_mask = checker ? checker->mask() : nullptr;
    golap_star_result r{};
    std::vector<int32_t> k(max_groups * std::max(1u, na));
    GroupResult out;
    out.sums.resize(max_groups);
    out.counts.resize(max_groups);
    detail::check(golap_star_run(dev.get(), &p, &r, k.data(), out.sums.data(), out.counts.data(), max_groups),
                  dev.get(), "star_query");
    out.sums.resize(r.n_groups);
    out.counts.resize(r.n_groups);
    for (uint64_t g = 0; g < r.n_groups; ++g) out.keys.emplace_back(k.begin() + g * na, k.begin() + (g + 1) * na);
    return out;
}

// The query 3 variants with one dimension's table on the CPU (03a.hpp, 03c.hpp, 03d.hpp):
// dims[host_dim] is joined through host_map (its map entry may be null), the others on the
// device; the groups equal star_query's.
inline GroupResult star_query_hostdim(Device &dev, const std::vector<StarDim> &dims, uint32_t host_dim,
                                      const HostHashMap &host_map, const CompInfo &value,
                                      const MetaChecker *checker = nullptr, uint64_t max_groups = 1 << 20,
                                      uint32_t threads = 0) {
    golap_star_plan p{};
    p.n_dims = (uint32_t)dims.size();
    uint32_t na = 0;
    for (size_t i = 0; i < dims.size(); ++i) {
        p.dims[i] = golap_star_dim{dims[i].fk->col.get(), dims[i].map ? dims[i].map->get() : nullptr,
                                   dims[i].group_lo, dims[i].group_card};
        na += dims[i].group_card ? 1 : 0;
    }
    p.value = value.col.get();
    p.chunk_end = value.blocks.size();
    p.dev_chunk_mask = checker ? checker->mask() : nullptr;
    golap_star_result r{};
    std::vector<int32_t> k(max_groups * std::max(1u, na));
    GroupResult out;
    out.sums.resize(max_groups);
    out.counts.resize(max_groups);
    detail::check(golap_star_run_hostdim(dev.get(), &p, host_dim, host_map.get(), threads, &r, nullptr, k.data(),
                                         out.sums.data(), out.counts.data(), max_groups),
                  dev.get(), "star_query_hostdim");
    out.sums.resize(r.n_groups);
    out.counts.resize(r.n_groups);
    for (uint64_t g = 0; g < r.n_groups; ++g) out.keys.emplace_back(k.begin() + g * na, k.begin() + (g + 1) * na);
    return out;
}

// query1 from SSD (01.hpp:350-424): the four columns' files (orderdate, discount, quantity,
// extendedprice) streamed in windows of chunks, pruned chunks never read; the same result as
// query1 over the resident columns.
inline Q1Result query1_files(Device &dev, const std::array<std::string, 4> &paths, const HashMap &date_map,
                             int64_t disc_lo, int64_t disc_hi, int64_t qty_lo, int64_t qty_hi,
                             int64_t price_lo = INT32_MIN, int64_t price_hi = INT32_MAX, uint32_t window_chunks = 0,
                             golap_stream_stats *stats = nullptr) {
    golap_q1_file_plan p{};
    for (int i = 0; i < 4; ++i) p.paths[i] = paths[i].c_str();
    p.date_map = date_map.get();
    p.disc_lo = disc_lo;
    p.disc_hi = disc_hi;
    p.qty_lo = qty_lo;
    p.qty_hi = qty_hi;
    p.price_lo = price_lo;
    p.price_hi = price_hi;
    p.window_chunks = window_chunks;
    p.zone_prune = 1;
    golap_q1_result r{};
    detail::check(golap_q1_run_files(dev.get(), &p, &r, stats), dev.get(), "query1_files");
    return Q1Result{r.sum, r.matches, r.chunks_run, r.chunks_pruned, r.comp_bytes, r.kernel_ms};
}

}  // namespace golap
