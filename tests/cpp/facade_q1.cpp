// The C++ facade used the way the reference's plan code uses its operator API
// (usecases/ssb/queries/01.hpp:53-449): prepare_compressed_device per column with MinMaxMeta,
// HashMap + hash_map_build_pred_other on the date dimension, MetaChecker pruning, the fused
// query1, and a TableLoader<DecompressEnv> stepping every chunk (the select path,
// core/hl/select.cuh:11-83), the same loader stepped on worker streams with events into a
// HashAggregate, load<DecompressEnv> with and without pruned chunks; also a column coded on the device, a star plan with one dimension
// joined through a HostHashMap, and query1 streamed from column files. Results are checked
// against the oracle (oracle/liboracle.so).
// Prints "OK" and exits 0 on success.
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/golap/golap.hpp"
#include "../../oracle/oracle.h"

#define REQUIRE(c)                                                        \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);      \
            return 1;                                                     \
        }                                                                 \
    } while (0)

int main() {
    const uint64_t n = 3'000'000;
    std::vector<int32_t> d_key(2556), date(n), price(n);
    std::vector<int16_t> d_year(2556);
    std::vector<uint8_t> d_month(2556), disc(n), qty(n);
    golap_gen_date_dim(2556, d_key.data(), d_year.data(), d_month.data());
    golap_gen_q1_fact(7, 0, n, d_key.data(), 2556, date.data(), disc.data(), qty.data(), price.data(), 0);
    // sort by date so zone maps prune (SortHelper's dimension sort, usecases/ssb/helper.hpp)
    std::vector<uint32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return date[a] < date[b]; });
    auto permute = [&](auto &v) {
        auto c = v;
        for (uint64_t i = 0; i < n; ++i) v[i] = c[ord[i]];
    };
    permute(date);
    permute(disc);
    permute(qty);
    permute(price);

    golap::Device dev(0);
    const uint64_t tuples_per_chunk = (1 << 20) / sizeof(int32_t);  // 01.hpp:62-63
    golap::CompInfo ci_date, ci_disc, ci_qty, ci_price;
    for (auto *ci : {&ci_date, &ci_price}) ci->chunk_bytes = tuples_per_chunk * 4;
    for (auto *ci : {&ci_disc, &ci_qty}) ci->chunk_bytes = tuples_per_chunk * 1;
    for (auto *ci : {&ci_date, &ci_disc, &ci_qty, &ci_price}) ci->lz4_min_match = 16;
    golap::MinMaxMeta<int32_t> mm_date;
    golap::MinMaxMeta<uint8_t> mm_qty;
    golap::prepare_compressed_device(dev, date.data(), n, ci_date, &mm_date);
    golap::BloomMeta<uint8_t> bloom_disc(0.01, 256);
    golap::EqHistogram<uint8_t> hist_qty(16);
    golap::prepare_compressed_device(dev, disc.data(), n, ci_disc, nullptr, nullptr, &bloom_disc);
    golap::prepare_compressed_device(dev, qty.data(), n, ci_qty, &mm_qty, &hist_qty);
    golap::prepare_compressed_device(dev, price.data(), n, ci_price);
    const uint64_t chunk_num = ci_date.blocks.size();
    REQUIRE(chunk_num == (n + tuples_per_chunk - 1) / tuples_per_chunk);  // apps_tests.cu:12-65
    REQUIRE(ci_disc.blocks.size() == chunk_num);

    golap::DeviceMem dk(dev, golap::Tag<int32_t>{}, 2556), dy(dev, golap::Tag<int16_t>{}, 2556);
    dk.copy_from(d_key.data(), 2556);
    dy.copy_from(d_year.data(), 2556);
    golap::HashMap date_map(dev, 2556, dk.ptr<int32_t>());
    golap::hash_map_build_pred_other(date_map, 2556, golap::ConstPred<int16_t>(dy.ptr<int16_t>(), 1995));
    REQUIRE(date_map.filled() == 365);

    golap::MetaChecker checker(dev, chunk_num);
    checker.add_minmax<int32_t>(ci_date, 19950101, 19951231);
    checker.add_minmax<uint8_t>(ci_qty, 30, 50);
    auto mask = checker.to_host();
    uint64_t keep = 0;
    for (uint64_t c = 0; c < chunk_num; ++c) {
        const bool want = mm_date.check_pred_host(c, 19950101, 19951231) && mm_qty.check_pred_host(c, 30, 50);
        REQUIRE((mask[c] != 0) == want);
        keep += want;
    }
    REQUIRE(keep > 0 && keep < chunk_num);

    auto full = golap::query1(dev, ci_date, ci_disc, ci_qty, ci_price, date_map, 4, 6, 30, 50);
    auto pruned = golap::query1(dev, ci_date, ci_disc, ci_qty, ci_price, date_map, 4, 6, 30, 50, &checker);
    oracle_pred pred{d_year.data(), 3, 1995, 1995};
    oracle_q1_consts qc{4, 6, 30, 50, INT32_MIN, INT32_MAX};
    uint64_t want_matches = 0;
    const int64_t want = oracle_q1_plan(d_key.data(), 2556, &pred, 1, date.data(), disc.data(), qty.data(),
                                        price.data(), n, &qc, 4, &want_matches);
    REQUIRE(full.sum == want && full.matches == want_matches);
    REQUIRE(pruned.sum == want && pruned.matches == want_matches);
    REQUIRE(pruned.chunks_pruned == chunk_num - keep);

    // histogram and Bloom metadata (metadata.cuh:225-537, MetaChecker::add_hist / add_bloom)
    golap::MetaChecker hchecker(dev, chunk_num);
    hchecker.add_hist<uint8_t>(hist_qty, 30, 50);
    hchecker.add_bloom<uint8_t>(bloom_disc, 5);
    auto hmask = hchecker.to_host();
    {
        std::vector<int64_t> q64(qty.begin(), qty.end()), d64(disc.begin(), disc.end()), mins(chunk_num),
            maxs(chunk_num);
        std::vector<uint64_t> hist(chunk_num * 16);
        std::vector<uint16_t> f1(chunk_num), f2(chunk_num);
        oracle_eqhist_build(q64.data(), n, tuples_per_chunk, 16, mins.data(), maxs.data(), hist.data());
        oracle_eqhist_check(mins.data(), maxs.data(), hist.data(), chunk_num, 16, 30, 50, f1.data());
        uint64_t bn = 0, bk = 0;
        oracle_bloom_params(0.01, 256, &bn, &bk);
        std::vector<uint32_t> state(chunk_num * 256 / 4);
        oracle_bloom_build(d64.data(), n, tuples_per_chunk, 256, bk, state.data());
        oracle_bloom_check(state.data(), chunk_num, 256, bk, 5, f2.data());
        for (uint64_t c = 0; c < chunk_num; ++c) REQUIRE((hmask[c] != 0) == (f1[c] && f2[c]));
    }
    auto hpruned = golap::query1(dev, ci_date, ci_disc, ci_qty, ci_price, date_map, 4, 6, 30, 50, &hchecker);
    REQUIRE(hpruned.sum == want && hpruned.matches == want_matches);

    // join-key pruning (01.hpp:246-270): the date map's keys against the orderdate zone maps
    golap::MetaChecker jchecker(dev, chunk_num);
    jchecker.add_join_keys(date_map, ci_date);
    auto jmask = jchecker.to_host();
    uint64_t jkeep = 0;
    for (uint64_t c = 0; c < chunk_num; ++c) {
        bool want = false;  // some 1995 date key inside the chunk's [min, max]
        for (int k = 0; k < 2556 && !want; ++k)
            want = d_year[k] == 1995 && mm_date.check_pred_host(c, d_key[k], d_key[k]);
        REQUIRE((jmask[c] != 0) == want);
        jkeep += want;
    }
    auto jpruned = golap::query1(dev, ci_date, ci_disc, ci_qty, ci_price, date_map, 4, 6, 30, 50, &jchecker);
    REQUIRE(jpruned.sum == want && jpruned.matches == want_matches);
    REQUIRE(jpruned.chunks_pruned == chunk_num - jkeep);

    // values already in HBM, coded on the device (prepare_compressed_device over a device
    // Column): the same blocks, zone maps and query result as the host-coded column
    golap::DeviceMem dvals(dev, golap::Tag<int32_t>{}, n);
    dvals.copy_from(date.data(), n);
    golap::CompInfo ci_date_dev;
    ci_date_dev.chunk_bytes = tuples_per_chunk * 4;
    ci_date_dev.lz4_min_match = 16;
    golap::MinMaxMeta<int32_t> mm_date_dev;
    golap::prepare_compressed_device<int32_t>(dev, dvals, n, ci_date_dev, &mm_date_dev);
    REQUIRE(ci_date_dev.blocks.size() == chunk_num);
    for (uint64_t c = 0; c < chunk_num; ++c)
        REQUIRE(ci_date_dev.blocks[c].offset == ci_date.blocks[c].offset &&
                ci_date_dev.blocks[c].payload == ci_date.blocks[c].payload);
    REQUIRE(mm_date_dev.mins == mm_date.mins && mm_date_dev.maxs == mm_date.maxs);
    auto dq = golap::query1(dev, ci_date_dev, ci_disc, ci_qty, ci_price, date_map, 4, 6, 30, 50);
    REQUIRE(dq.sum == want && dq.matches == want_matches);

    // one dimension joined on the host (the query 3 variants with a table on the CPU): group by
    // year through a device map and by month through a HostHashMap over the date keys
    std::vector<int32_t> year32(d_year.begin(), d_year.end()), month32(d_month.begin(), d_month.end());
    golap::DeviceMem dyi(dev, golap::Tag<int32_t>{}, 2556), dmi(dev, golap::Tag<int32_t>{}, 2556);
    dyi.copy_from(year32.data(), 2556);
    dmi.copy_from(month32.data(), 2556);
    golap::HashMap ymap(dev, 2556, dk.ptr<int32_t>()), mmap(dev, 2556, dk.ptr<int32_t>());
    golap::hash_map_build_pred_payload(ymap, 2556, dyi.ptr<int32_t>());
    golap::hash_map_build_pred_payload(mmap, 2556, dmi.ptr<int32_t>());
    golap::HostHashMap hmonth(2556, d_key.data());
    hmonth.build(2556, month32.data());
    REQUIRE(hmonth.filled() == 2556 && hmonth.probe(d_key[100]) == (uint64_t)month32[100]);
    REQUIRE(hmonth.probe(-5) == UINT64_MAX);
    std::vector<golap::StarDim> sd = {{&ci_date, &ymap, 1992, 7}, {&ci_date, &mmap, 1, 12}};
    auto g_dev = golap::star_query(dev, sd, ci_price);
    sd[1].map = nullptr;
    auto g_host = golap::star_query_hostdim(dev, sd, 1, hmonth, ci_price);
    REQUIRE(!g_dev.sums.empty() && g_dev.keys == g_host.keys && g_dev.sums == g_host.sums &&
            g_dev.counts == g_host.counts);

    // query1 streamed from column files (query1 from SSD, 01.hpp:350-424)
    {
        golap_encode_opts o;
        golap_encode_opts_default(&o);
        o.codec = GOLAP_CODEC_LZ4;
        o.lz4_min_match = 16;
        std::array<std::string, 4> paths;
        const void *vals[4] = {date.data(), disc.data(), qty.data(), price.data()};
        const uint32_t types[4] = {GOLAP_I32, GOLAP_U8, GOLAP_U8, GOLAP_I32};
        for (int i = 0; i < 4; ++i) {
            o.chunk_rows = (uint32_t)tuples_per_chunk;
            golap_host_column *h = nullptr;
            REQUIRE(golap_encode(vals[i], n, types[i], &o, &h) == GOLAP_OK);
            paths[i] = "/tmp/golap_facade_" + std::to_string(getpid()) + "_" + std::to_string(i) + ".gcol";
            REQUIRE(golap_host_column_save(h, paths[i].c_str()) == GOLAP_OK);
            golap_host_column_free(h);
        }
        golap_stream_stats st{};
        auto fq = golap::query1_files(dev, paths, date_map, 4, 6, 30, 50, INT32_MIN, INT32_MAX, 3, &st);
        for (auto &p : paths) std::remove(p.c_str());
        REQUIRE(fq.sum == want && fq.matches == want_matches);
        REQUIRE(st.windows + st.windows_skipped == (chunk_num + 2) / 3);
    }

    // select path: step every chunk of every column through a TableLoader
    std::vector<uint64_t> all(chunk_num);
    std::iota(all.begin(), all.end(), 0);
    golap::TableLoader<golap::DecompressEnv> loader(dev);
    loader.add("lo_orderdate", all, 0, chunk_num, ci_date);
    loader.add("lo_discount", all, 0, chunk_num, ci_disc);
    std::vector<int32_t> chunk_d(tuples_per_chunk);
    std::vector<uint8_t> chunk_s(tuples_per_chunk);
    uint64_t c = 0;
    while (loader.rootop.step()) {
        const uint64_t rows = ci_date.blocks[c].tuples;
        loader.blockenvs[0]->decomp_buf().copy_to(chunk_d.data(), tuples_per_chunk);
        loader.blockenvs[1]->decomp_buf().copy_to(chunk_s.data(), tuples_per_chunk);
        REQUIRE(!std::memcmp(chunk_d.data(), date.data() + c * tuples_per_chunk, rows * 4));
        REQUIRE(!std::memcmp(chunk_s.data(), disc.data() + c * tuples_per_chunk, rows));
        REQUIRE(loader.blockenvs[0]->root.last_produced == rows * 4);
        ++c;
    }
    REQUIRE(c == chunk_num);

    // the stream protocol (core/core.hpp:94-124, 252-307): a Collector steps the discount,
    // quantity and price columns on their envs' streams behind the caller's event, the caller's
    // stream waits for the collector's event, and a HashAggregate adds every chunk's rows
    // (group by discount, quantity; SUM(price)) on that stream; checked against the oracle
    {
        golap::TableLoader<golap::DecompressEnv> tl(dev);
        tl.add("lo_discount", all, 0, chunk_num, ci_disc);
        tl.add("lo_quantity", all, 0, chunk_num, ci_qty);
        tl.add("lo_price", all, 0, chunk_num, ci_price);
        std::vector<golap_group_key> gk(2);
        gk[0].type = GOLAP_U8, gk[0].lo = 0, gk[0].card = 11;
        gk[1].type = GOLAP_U8, gk[1].lo = 1, gk[1].card = 50;
        golap::HashAggregate agg(dev, gk);
        uint64_t steps = 0;
        tl.rootevent.record(tl.rootstream.stream);
        while (tl.rootop.step(tl.rootstream.stream, tl.rootevent.event)) {
            const uint64_t rows = ci_disc.blocks[steps].tuples;
            agg.add<int32_t>({tl.at("lo_discount").decomp_buf().ptr<void>(), tl.at("lo_quantity").decomp_buf().ptr<void>()},
                             tl.at("lo_price").decomp_buf().ptr<int32_t>(), rows, tl.rootstream.stream);
            // the next step's decode into the same buffers waits for this chunk's adds
            tl.rootevent.record(tl.rootstream.stream);
            ++steps;
        }
        REQUIRE(steps == chunk_num);
        tl.rootstream.sync();
        auto got = agg.results(11 * 50);
        std::vector<int32_t> k0(disc.begin(), disc.end()), k1(qty.begin(), qty.end());
        std::vector<int64_t> v64(price.begin(), price.end());
        const int32_t *keys[2] = {k0.data(), k1.data()};
        std::vector<int32_t> wk(550 * 2);
        std::vector<int64_t> ws(550);
        std::vector<uint64_t> wc(550);
        const int64_t ng = oracle_groupby(keys, 2, v64.data(), n, 550, wk.data(), ws.data(), wc.data());
        REQUIRE(ng > 0 && (uint64_t)ng == got.sums.size());
        for (int64_t g = 0; g < ng; ++g) {
            REQUIRE(got.keys[g][0] == wk[2 * g] && got.keys[g][1] == wk[2 * g + 1]);
            REQUIRE(got.sums[g] == ws[g] && got.counts[g] == wc[g]);
        }
        // a second pass accumulates on top; clear() starts over
        agg.clear();
        REQUIRE(agg.results(550).sums.empty());
    }

    // load<BLOCKENV, T> (core/apps.cuh:682-793): three workers decode the date column into host
    // memory, once whole and once with the MetaChecker's pruned chunks skipped
    {
        std::vector<int32_t> host(chunk_num * tuples_per_chunk, -1);
        uint64_t alloc = 0, skipped = 0;
        const double ms = golap::load<golap::DecompressEnv>(dev, 3, host.data(), ci_date, alloc);
        REQUIRE(ms >= 0 && alloc > 0 && !std::memcmp(host.data(), date.data(), n * 4));
        std::vector<int32_t> host2(chunk_num * tuples_per_chunk, -1);
        golap::load<golap::DecompressEnv>(dev, 3, host2.data(), ci_date, alloc, mask.data(), &skipped);
        REQUIRE(skipped == chunk_num - keep);
        for (uint64_t c2 = 0; c2 < chunk_num; ++c2) {
            const uint64_t rows = ci_date.blocks[c2].tuples;
            const int32_t *got = host2.data() + c2 * tuples_per_chunk;
            if (mask[c2]) {
                REQUIRE(!std::memcmp(got, date.data() + c2 * tuples_per_chunk, rows * 4));
            } else {
                REQUIRE(std::all_of(got, got + rows, [](int32_t v) { return v == -1; }));
            }
        }
    }
    std::printf("OK sum=%lld matches=%llu chunks=%llu pruned=%llu\n", (long long)full.sum,
                (unsigned long long)full.matches, (unsigned long long)chunk_num,
                (unsigned long long)pruned.chunks_pruned);
    return 0;
}
