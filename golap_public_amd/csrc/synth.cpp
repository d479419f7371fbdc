// Seeded synthetic columns for the benchmark workloads (host, multithreaded).
//
// Counter-based: value(seed, column, row) = lo + (((splitmix64(seed*PHI ^ col<<48 ^ row) >> 32)
// * span) >> 32), the multiply-shift range reduction of util::uniform_int (core/util.hpp:177-181)
// applied to a SplitMix64 counter hash, so any row range can be generated independently.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/golap_hip.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint32_t uniform(uint64_t seed, uint32_t col, uint64_t row, uint32_t lo, uint32_t hi) {
    const uint64_t r = splitmix64((seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)col << 48) ^ row);
    const uint64_t span = (uint64_t)hi - lo + 1;
    return lo + (uint32_t)(((r >> 32) * span) >> 32);
}

// 53 random bits of (seed, column, row).
inline uint64_t bits53(uint64_t seed, uint32_t col, uint64_t row) {
    return splitmix64((seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)col << 48) ^ row) >> 11;
}

// exp(mu + sigma * z) for a standard normal z drawn by Box-Muller from two columns, rounded
// half up and clipped to [lo, hi].
inline int32_t lognormal(uint64_t seed, uint32_t col, uint64_t row, double mu, double sigma, int64_t lo,
                         int64_t hi) {
    const double u1 = (double)(bits53(seed, col, row) + 1) * 0x1p-53;  // (0, 1]
    const double u2 = (double)bits53(seed, col + 1, row) * 0x1p-53;     // [0, 1)
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    const double x = std::floor(std::exp(mu + sigma * z) + 0.5);
    const double c = std::min((double)hi, std::max((double)lo, x));
    return (int32_t)(int64_t)c;
}

template <typename F>
void parallel_rows(uint64_t n, uint32_t threads, F f) {
    if (!threads) threads = std::max(1u, std::thread::hardware_concurrency());
    threads = (uint32_t)std::min<uint64_t>(threads, std::max<uint64_t>(1, n / 65536));
    std::vector<std::thread> ts;
    for (uint32_t t = 0; t < threads; ++t)
        ts.emplace_back([&, t] { f(n * t / threads, n * (t + 1) / threads); });
    for (auto &th : ts) th.join();
}

// Civil date from days since 1970-01-01 (H. Hinnant, "chrono-Compatible Low-Level Date
// Algorithms", civil_from_days).
void civil_from_days(int64_t z, int &y, unsigned &m, unsigned &d) {
    z += 719468;
    const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
    const unsigned doe = (unsigned)(z - era * 146097);
    const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
    const int64_t yy = (int64_t)yoe + era * 400;
    const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
    const unsigned mp = (5 * doy + 2) / 153;
    d = doy - (153 * mp + 2) / 5 + 1;
    m = mp < 10 ? mp + 3 : mp - 9;
    y = (int)(yy + (m <= 2));
}

}  // namespace

extern "C" {

int golap_gen_date_dim(uint32_t n_days, int32_t *key, int16_t *year, uint8_t *month) {
    for (uint32_t i = 0; i < n_days; ++i) {
        int y;
        unsigned m, d;
        civil_from_days(8035 + (int64_t)i, y, m, d);  // day 8035 = 1992-01-01
        if (key) key[i] = y * 10000 + (int)m * 100 + (int)d;
        if (year) year[i] = (int16_t)y;
        if (month) month[i] = (uint8_t)m;
    }
    return GOLAP_OK;
}

int golap_gen_q1_fact(uint64_t seed, uint64_t row0, uint64_t n, const int32_t *date_keys,
                      uint32_t n_dates, int32_t *date_key, uint8_t *discount, uint8_t *quantity,
                      int32_t *price, uint32_t threads) {
    if ((date_key && (!date_keys || !n_dates))) return GOLAP_EINVAL;
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t r = row0 + i;
            if (date_key) date_key[i] = date_keys[uniform(seed, 0, r, 0, n_dates - 1)];
            if (discount) discount[i] = (uint8_t)uniform(seed, 1, r, 0, 10);
            if (quantity) quantity[i] = (uint8_t)uniform(seed, 2, r, 1, 50);
            if (price) price[i] = (int32_t)uniform(seed, 3, r, 90000, 10500000);
        }
    });
    return GOLAP_OK;
}

}  // extern "C"

extern "C" {

// Star schema fact rows [row0, row0+n) of n_total: date_key sorted (row r falls on day
// floor(r * n_dates / n_total)), cust/supp/part keys uniform over the dimension sizes, revenue
// U[100, 5000000]. Any output may be NULL.
int golap_gen_star_fact(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total,
                        const int32_t *date_keys, uint32_t n_dates, uint32_t n_cust,
                        uint32_t n_supp, uint32_t n_part, int32_t *date_key, int32_t *cust_key,
                        int32_t *supp_key, int32_t *part_key, int32_t *revenue, uint32_t threads) {
    if ((date_key && (!date_keys || !n_dates || !n_total)) || (cust_key && !n_cust) ||
        (supp_key && !n_supp) || (part_key && !n_part))
        return GOLAP_EINVAL;
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t r = row0 + i;
            if (date_key) date_key[i] = date_keys[(unsigned __int128)r * n_dates / n_total];
            if (cust_key) cust_key[i] = (int32_t)uniform(seed, 10, r, 0, n_cust - 1);
            if (supp_key) supp_key[i] = (int32_t)uniform(seed, 11, r, 0, n_supp - 1);
            if (part_key) part_key[i] = (int32_t)uniform(seed, 12, r, 0, n_part - 1);
            if (revenue) revenue[i] = (int32_t)uniform(seed, 13, r, 100, 5000000);
        }
    });
    return GOLAP_OK;
}

// Trip-table rows [row0, row0 + n) of n_total.
int golap_gen_trip(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total, int64_t *pickup_ts,
                   uint8_t *passenger, uint16_t *zone, int32_t *fare, int32_t *distance, uint32_t threads) {
    if (pickup_ts && !n_total) return GOLAP_EINVAL;
    // Zipf(s = 1.2) over 1..6 by inverse CDF on 53 bits: thresholds floor(cdf_k * 2^53)
    double w[6], tot = 0;
    for (int k = 0; k < 6; ++k) tot += (w[k] = std::pow((double)(k + 1), -1.2));
    uint64_t thr[5];
    double cum = 0;
    for (int k = 0; k < 5; ++k) {
        cum += w[k];
        thr[k] = (uint64_t)std::floor(cum / tot * 0x1p53);
    }
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t r = row0 + i;
            if (pickup_ts)
                pickup_ts[i] = GOLAP_TRIP_T0 + (int64_t)((unsigned __int128)r * GOLAP_TRIP_SPAN / n_total);
            if (passenger) {
                const uint64_t u = bits53(seed, 30, r);
                uint8_t p = 1;
                for (int k = 0; k < 5; ++k) p += u >= thr[k];
                passenger[i] = p;
            }
            if (zone) zone[i] = (uint16_t)uniform(seed, 31, r, 1, 265);
            if (fare) fare[i] = lognormal(seed, 32, r, 7.2, 0.6, 250, 50000);
            if (distance) distance[i] = lognormal(seed, 34, r, 8.0, 0.8, 1, INT32_MAX);
        }
    });
    return GOLAP_OK;
}

// A dimension attribute uniform over [lo, hi] for rows [0, n) (column id `col`).
int golap_gen_uniform_i32(uint64_t seed, uint32_t col, uint64_t n, int32_t lo, int32_t hi,
                          int32_t *out, uint32_t threads) {
    if (!out || hi < lo) return GOLAP_EINVAL;
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i)
            out[i] = lo + (int32_t)uniform(seed, col, i, 0, (uint32_t)((int64_t)hi - lo));
    });
    return GOLAP_OK;
}

}  // extern "C"
