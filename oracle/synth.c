/*
 * Seeded synthetic columns (TEST INFRASTRUCTURE: the oracle, never the product).
 *
 * Counter-based so that any row range can be produced independently and in parallel:
 * r = splitmix64(seed*PHI ^ (col << 48) ^ row), value = lo + (((r >> 32) * span) >> 32), the
 * same multiply-shift range reduction as util::uniform_int (core/util.hpp:177-181) applied to a
 * counter hash instead of an LCG stream.
 */
#include <math.h>

#include "oracle.h"

/* SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
 * OOPSLA 2014). */
uint64_t oracle_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint32_t oracle_uniform(uint64_t seed, uint32_t col, uint64_t row, uint32_t lo, uint32_t hi) {
    uint64_t r = oracle_splitmix64((seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)col << 48) ^ row);
    uint64_t span = (uint64_t)hi - lo + 1;
    return lo + (uint32_t)(((r >> 32) * span) >> 32);
}

void oracle_gen_q1_fact(uint64_t seed, uint64_t row0, uint64_t n, const int32_t *date_keys,
                        uint32_t n_dates, int32_t *date_key, uint8_t *discount,
                        uint8_t *quantity, int32_t *price) {
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t r = row0 + i;
        if (date_key) date_key[i] = date_keys[oracle_uniform(seed, 0, r, 0, n_dates - 1)];
        if (discount) discount[i] = (uint8_t)oracle_uniform(seed, 1, r, 0, 10);
        if (quantity) quantity[i] = (uint8_t)oracle_uniform(seed, 2, r, 1, 50);
        if (price) price[i] = (int32_t)oracle_uniform(seed, 3, r, 90000, 10500000);
    }
}

/* Civil date from a day count since 1970-01-01 (H. Hinnant, "chrono-Compatible Low-Level Date
 * Algorithms", civil_from_days). */
static void civil_from_days(int64_t z, int *y, unsigned *m, unsigned *d) {
    z += 719468;
    int64_t era = (z >= 0 ? z : z - 146096) / 146097;
    unsigned doe = (unsigned)(z - era * 146097);
    unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
    int64_t yy = (int64_t)yoe + era * 400;
    unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
    unsigned mp = (5 * doy + 2) / 153;
    *d = doy - (153 * mp + 2) / 5 + 1;
    *m = mp < 10 ? mp + 3 : mp - 9;
    *y = (int)(yy + (*m <= 2));
}

void oracle_gen_date_dim(uint32_t n_days, int32_t *key, int16_t *year, uint8_t *month) {
    const int64_t day0 = 8035; /* 1992-01-01 */
    for (uint32_t i = 0; i < n_days; ++i) {
        int y;
        unsigned m, d;
        civil_from_days(day0 + i, &y, &m, &d);
        if (key) key[i] = y * 10000 + (int)m * 100 + (int)d;
        if (year) year[i] = (int16_t)y;
        if (month) month[i] = (uint8_t)m;
    }
}

/* Star-schema fact columns, the same formulas as the product generator (restated here so the
 * checker does not depend on the product). */
void oracle_gen_star_fact(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total,
                          const int32_t *date_keys, uint32_t n_dates, uint32_t n_cust,
                          uint32_t n_supp, uint32_t n_part, int32_t *date_key, int32_t *cust_key,
                          int32_t *supp_key, int32_t *part_key, int32_t *revenue) {
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t r = row0 + i;
        if (date_key) date_key[i] = date_keys[(unsigned __int128)r * n_dates / n_total];
        if (cust_key) cust_key[i] = (int32_t)oracle_uniform(seed, 10, r, 0, n_cust - 1);
        if (supp_key) supp_key[i] = (int32_t)oracle_uniform(seed, 11, r, 0, n_supp - 1);
        if (part_key) part_key[i] = (int32_t)oracle_uniform(seed, 12, r, 0, n_part - 1);
        if (revenue) revenue[i] = (int32_t)oracle_uniform(seed, 13, r, 100, 5000000);
    }
}

/* Trip-table columns: the same definitions as the product generator, restated. */
static uint64_t bits53(uint64_t seed, uint32_t col, uint64_t row) {
    return oracle_splitmix64((seed * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)col << 48) ^ row) >> 11;
}

static int32_t lognormal(uint64_t seed, uint32_t col, uint64_t row, double mu, double sigma, double lo,
                         double hi) {
    double u1 = (double)(bits53(seed, col, row) + 1) * 0x1p-53;
    double u2 = (double)bits53(seed, col + 1, row) * 0x1p-53;
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double x = floor(exp(mu + sigma * z) + 0.5);
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return (int32_t)(int64_t)x;
}

void oracle_gen_trip(uint64_t seed, uint64_t row0, uint64_t n, uint64_t n_total, int64_t *pickup_ts,
                     uint8_t *passenger, uint16_t *zone, int32_t *fare, int32_t *distance) {
    double w[6], tot = 0, cum = 0;
    uint64_t thr[5];
    for (int k = 0; k < 6; ++k) {
        w[k] = pow((double)(k + 1), -1.2);
        tot += w[k];
    }
    for (int k = 0; k < 5; ++k) {
        cum += w[k];
        thr[k] = (uint64_t)floor(cum / tot * 0x1p53);
    }
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t r = row0 + i;
        if (pickup_ts)
            pickup_ts[i] = 1704067200LL + (int64_t)((unsigned __int128)r * 31622400u / n_total);
        if (passenger) {
            uint64_t u = bits53(seed, 30, r);
            uint8_t p = 1;
            for (int k = 0; k < 5; ++k) p += u >= thr[k];
            passenger[i] = p;
        }
        if (zone) zone[i] = (uint16_t)oracle_uniform(seed, 31, r, 1, 265);
        if (fare) fare[i] = lognormal(seed, 32, r, 7.2, 0.6, 250, 50000);
        if (distance) distance[i] = lognormal(seed, 34, r, 8.0, 0.8, 1, 2147483647.0);
    }
}

void oracle_gen_uniform_i32(uint64_t seed, uint32_t col, uint64_t n, int32_t lo, int32_t hi,
                            int32_t *out) {
    for (uint64_t i = 0; i < n; ++i)
        out[i] = lo + (int32_t)oracle_uniform(seed, col, i, 0, (uint32_t)((int64_t)hi - lo));
}
