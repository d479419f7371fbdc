/*
 * SSB data generator restatement (TEST INFRASTRUCTURE: the oracle, never the product).
 *
 * Follows SSB_Tables_col (usecases/ssb/data/SSB_def.hpp:141-423) call for call, because the
 * single LCG state is threaded through every table in populate() order
 * (date, customer, supplier, part, lineorder; SSB_def.hpp:179-185). The reference's logged
 * results (usecases/ssb/results/query_sf{10,50,100}_export.csv) are only reproducible if every
 * uniform_int call happens in the same order with the same bounds.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* util::uniform_int, core/util.hpp:177-181: LCG step, then multiply-shift into [a, b). */
static inline uint32_t lcg_uniform(uint64_t *seed, uint32_t a, uint32_t b) {
    *seed = *seed * 1103515245ull + 12345ull;
    return (uint32_t)((((*seed) >> 32) * (uint64_t)(b - a)) >> 32) + a;
}

/* SSB_helper.hpp:67-73 */
static inline uint32_t partkey_to_price(uint64_t p) {
    uint32_t price = 90000;
    price += (uint32_t)((p / 10) % 20001);
    price += (uint32_t)((p % 1000) * 100);
    return price;
}

/* NATION is a std::map (SSB_helper.hpp:39-65), so populate_supplier/customer index its keys in
 * sorted order; this is that order with each nation's region index (REGION, SSB_helper.hpp:38). */
static const int NATION_REGION_SORTED[25] = {
    0 /*ALGERIA*/,   1 /*ARGENTINA*/, 1 /*BRAZIL*/,     1 /*CANADA*/,   2 /*CHINA*/,
    4 /*EGYPT*/,     0 /*ETHIOPIA*/,  3 /*FRANCE*/,     3 /*GERMANY*/,  2 /*INDIA*/,
    2 /*INDONESIA*/, 4 /*IRAN*/,      4 /*IRAQ*/,       2 /*JAPAN*/,    4 /*JORDAN*/,
    0 /*KENYA*/,     0 /*MOROCCO*/,   0 /*MOZAMBIQUE*/, 1 /*PERU*/,     3 /*ROMANIA*/,
    3 /*RUSSIA*/,    4 /*SAUDI ARABIA*/, 3 /*UNITED KINGDOM*/, 1 /*UNITED STATES*/,
    2 /*VIETNAM*/};

#define SSB_DATES 2555u          /* 365*7, SSB_def.hpp:151 */
#define CUST_MORTALITY 3u        /* SSB_helper.hpp:8 */
#define N_COLOR 92u              /* COLOR.size(), SSB_helper.hpp:12-21 */
#define N_PRIORITY 5u
#define N_SHIPMODE 7u
#define N_SEGMENT 5u

void oracle_ssb_sizes(uint32_t sf, uint64_t *orders, uint64_t *parts, uint64_t *suppliers,
                      uint64_t *customers, uint64_t *dates) {
    if (orders) *orders = (uint64_t)sf * 1500000ull;
    if (parts) *parts = (uint64_t)(200000.0 * floor(1.0 + log2((double)sf)));
    if (suppliers) *suppliers = (uint64_t)sf * 2000ull;
    if (customers) *customers = (uint64_t)sf * 30000ull; /* customer_factor = 1, ssb.cu:81 */
    if (dates) *dates = SSB_DATES;
}

typedef struct {
    int32_t key[SSB_DATES];
    int16_t year[SSB_DATES];
    int32_t yearmonthnum[SSB_DATES];
    int16_t weeknuminyear[SSB_DATES];
} ssb_dates;

/* populate_date, SSB_def.hpp:372-399 (no LCG calls). */
static void gen_dates(ssb_dates *d) {
    time_t cur = 694245661; /* Wed Jan 1 06:01:01 1992 GMT */
    for (uint32_t i = 0; i < SSB_DATES; ++i) {
        struct tm tm;
        gmtime_r(&cur, &tm);
        d->key[i] = (1900 + tm.tm_year) * 10000 + (tm.tm_mon + 1) * 100 + tm.tm_mday;
        d->year[i] = (int16_t)(1900 + tm.tm_year);
        d->yearmonthnum[i] = (1900 + tm.tm_year) * 100 + tm.tm_mon + 1;
        d->weeknuminyear[i] =
            (int16_t)(((tm.tm_yday + 7 - (tm.tm_wday ? (tm.tm_wday - 1) : 6)) / 7) + 1);
        cur += 60 * 60 * 24;
    }
}

typedef struct {
    uint64_t seed;
    uint64_t orders, parts, suppliers, customers;
    ssb_dates dates;
    int16_t *p_category, *p_brand1;
    uint8_t *s_region;
} ssb_state;

/* populate_customer (SSB_def.hpp:344-370): 6 draws per row; only the LCG state matters for
 * the Q1/Q2 plans. */
static void gen_customers(ssb_state *s) {
    for (uint64_t i = 0; i < s->customers; ++i) {
        lcg_uniform(&s->seed, 0, 25);     /* nation_idx */
        lcg_uniform(&s->seed, 0, 10);     /* city_code */
        lcg_uniform(&s->seed, 100, 1000); /* acode */
        lcg_uniform(&s->seed, 100, 1000); /* exchg */
        lcg_uniform(&s->seed, 1000, 10000); /* pnumber */
        lcg_uniform(&s->seed, 0, N_SEGMENT); /* mktsegment */
    }
}

/* populate_supplier (SSB_def.hpp:318-342): 5 draws per row. */
static void gen_suppliers(ssb_state *s) {
    for (uint64_t i = 0; i < s->suppliers; ++i) {
        uint32_t nation = lcg_uniform(&s->seed, 0, 25);
        lcg_uniform(&s->seed, 0, 10);
        lcg_uniform(&s->seed, 100, 1000);
        lcg_uniform(&s->seed, 100, 1000);
        lcg_uniform(&s->seed, 1000, 10000);
        s->s_region[i] = (uint8_t)NATION_REGION_SORTED[nation];
    }
}

/* populate_part (SSB_def.hpp:281-316): 12 draws per row; category "MFGR#<num1><num2>",
 * brand1 "MFGR#<num1><num2><num3:02>" kept as integers num1*10+num2 and
 * num1*1000+num2*100+num3 (every brand1 string has the same length, so string order is
 * integer order). */
static void gen_parts(ssb_state *s) {
    for (uint64_t i = 0; i < s->parts; ++i) {
        lcg_uniform(&s->seed, 1, 51); /* p_size */
        uint32_t num1 = lcg_uniform(&s->seed, 1, 6);
        uint32_t num2 = lcg_uniform(&s->seed, 1, 6);
        uint32_t num3 = lcg_uniform(&s->seed, 1, 41);
        lcg_uniform(&s->seed, 0, 6);  /* tsyl1: TYPE_SYL1.size() */
        lcg_uniform(&s->seed, 0, 5);  /* tsyl2 */
        lcg_uniform(&s->seed, 0, 5);  /* tsyl3 */
        lcg_uniform(&s->seed, 0, 5);  /* csyl1: CONT_SYL1.size() */
        lcg_uniform(&s->seed, 0, 8);  /* csyl2 */
        lcg_uniform(&s->seed, 0, N_COLOR); /* name1 */
        lcg_uniform(&s->seed, 0, N_COLOR); /* name2 */
        lcg_uniform(&s->seed, 0, N_COLOR); /* col1 */
        s->p_category[i] = (int16_t)(num1 * 10 + num2);
        s->p_brand1[i] = (int16_t)(num1 * 1000 + num2 * 100 + num3);
    }
}

static int ssb_init(ssb_state *s, uint32_t sf) {
    memset(s, 0, sizeof(*s));
    s->seed = 0xBEE5BEE5ull; /* SSB_def.hpp:416 */
    uint64_t dates;
    oracle_ssb_sizes(sf, &s->orders, &s->parts, &s->suppliers, &s->customers, &dates);
    gen_dates(&s->dates);
    s->p_category = (int16_t *)malloc(s->parts * sizeof(int16_t));
    s->p_brand1 = (int16_t *)malloc(s->parts * sizeof(int16_t));
    s->s_region = (uint8_t *)malloc(s->suppliers);
    if (!s->p_category || !s->p_brand1 || !s->s_region) return -1;
    gen_customers(s);
    gen_suppliers(s);
    gen_parts(s);
    return 0;
}

static void ssb_free(ssb_state *s) {
    free(s->p_category);
    free(s->p_brand1);
    free(s->s_region);
}

typedef struct {
    int32_t orderdate;
    uint32_t date_idx;
    uint8_t quantity, discount;
    uint32_t extendedprice, revenue;
    uint32_t partkey, suppkey;
} ssb_line;

/* populate_lineorder (SSB_def.hpp:201-279). Calls fn for each generated line in order. */
typedef void (*line_fn)(void *ctx, uint64_t row, const ssb_line *l);
static uint64_t gen_lineorder(ssb_state *s, line_fn fn, void *ctx) {
    uint64_t total = 0;
    for (uint64_t o = 0; o < s->orders; ++o) {
        uint32_t lines = lcg_uniform(&s->seed, 1, 8);
        uint32_t didx = lcg_uniform(&s->seed, 0, SSB_DATES);
        uint32_t cust;
        do cust = lcg_uniform(&s->seed, 0, (uint32_t)s->customers);
        while (cust % CUST_MORTALITY == 0);
        lcg_uniform(&s->seed, 0, N_PRIORITY); /* order priority */
        for (uint32_t l = 0; l < lines; ++l) {
            ssb_line x;
            x.date_idx = didx;
            x.orderdate = s->dates.key[didx];
            x.quantity = (uint8_t)lcg_uniform(&s->seed, 1, 51);
            x.discount = (uint8_t)lcg_uniform(&s->seed, 0, 11);
            lcg_uniform(&s->seed, 0, 9); /* tax */
            x.partkey = lcg_uniform(&s->seed, 0, (uint32_t)s->parts);
            x.suppkey = lcg_uniform(&s->seed, 0, (uint32_t)s->suppliers);
            x.extendedprice = x.quantity * partkey_to_price(x.partkey);
            /* lo_revenue = cur_extendedprice * (100.0-cur_discount)/100 stored as uint32
             * (SSB_def.hpp:251): double arithmetic, truncated on assignment. */
            x.revenue = (uint32_t)((double)x.extendedprice * (100.0 - (double)x.discount) / 100.0);
            lcg_uniform(&s->seed, 30, 91);          /* commitdate offset */
            lcg_uniform(&s->seed, 0, N_SHIPMODE);   /* shipmode */
            if (fn) fn(ctx, total + l, &x);
        }
        total += lines;
    }
    return total;
}

/* ---- golden evaluation: Q1.x (01inmem.hpp:51-141), Q2.x (02inmem.hpp:58-206) ---- */
#define Q2_YEARS 7
#define Q2_BRANDS 6000
typedef struct {
    ssb_state *s;
    uint64_t q1[3];
    uint64_t *q2sum[3];
    uint64_t *q2cnt[3];
} golden_ctx;

static void golden_line(void *vctx, uint64_t row, const ssb_line *x) {
    (void)row;
    golden_ctx *g = (golden_ctx *)vctx;
    const ssb_dates *d = &g->s->dates;
    uint32_t di = x->date_idx;
    /* Q1.1: d_year = 1993, discount 1..3, quantity 0..24 (01.hpp:451-463) */
    if (d->year[di] == 1993 && x->discount >= 1 && x->discount <= 3 && x->quantity <= 24)
        g->q1[0] += (uint64_t)(uint32_t)(x->extendedprice * x->discount);
    /* Q1.2: d_yearmonthnum = 199401, discount 4..6, quantity 26..35 */
    if (d->yearmonthnum[di] == 199401 && x->discount >= 4 && x->discount <= 6 &&
        x->quantity >= 26 && x->quantity <= 35)
        g->q1[1] += (uint64_t)(uint32_t)(x->extendedprice * x->discount);
    /* Q1.3: d_weeknuminyear = 6 and d_year = 1994, discount 5..7, quantity 26..35 */
    if (d->weeknuminyear[di] == 6 && d->year[di] == 1994 && x->discount >= 5 &&
        x->discount <= 7 && x->quantity >= 26 && x->quantity <= 35)
        g->q1[2] += (uint64_t)(uint32_t)(x->extendedprice * x->discount);

    /* Q2.x: part predicate, supplier region, group (d_year, p_brand1), SUM(lo_revenue) */
    int16_t cat = g->s->p_category[x->partkey];
    int16_t brand = g->s->p_brand1[x->partkey];
    uint8_t reg = g->s->s_region[x->suppkey];
    int q = -1;
    if (cat == 12 && reg == 1) q = 0;                                   /* MFGR#12, AMERICA */
    else if (brand >= 2221 && brand <= 2228 && reg == 2) q = 1;         /* MFGR#2221..8, ASIA */
    else if (brand == 2239 && reg == 3) q = 2;                          /* MFGR#2239, EUROPE */
    if (q >= 0) {
        uint32_t slot = (uint32_t)(d->year[di] - 1992) * Q2_BRANDS + (uint32_t)brand;
        g->q2sum[q][slot] += x->revenue;
        g->q2cnt[q][slot] += 1;
    }
}

int oracle_ssb_golden_run(uint32_t sf, oracle_ssb_golden *out) {
    ssb_state s;
    if (ssb_init(&s, sf)) return -1;
    golden_ctx g;
    memset(&g, 0, sizeof(g));
    g.s = &s;
    for (int q = 0; q < 3; ++q) {
        g.q2sum[q] = (uint64_t *)calloc(Q2_YEARS * Q2_BRANDS, sizeof(uint64_t));
        g.q2cnt[q] = (uint64_t *)calloc(Q2_YEARS * Q2_BRANDS, sizeof(uint64_t));
    }
    out->lineorder_rows = gen_lineorder(&s, golden_line, &g);
    for (int q = 0; q < 3; ++q) {
        out->q1[q] = g.q1[q];
        uint64_t sum = 0, groups = 0;
        for (uint32_t i = 0; i < Q2_YEARS * Q2_BRANDS; ++i) {
            if (g.q2cnt[q][i]) {
                groups++;
                sum += g.q2sum[q][i];
            }
        }
        out->q2_sum[q] = sum;
        out->q2_groups[q] = groups;
        free(g.q2sum[q]);
        free(g.q2cnt[q]);
    }
    ssb_free(&s);
    return 0;
}

typedef struct {
    int32_t *orderdate;
    uint8_t *discount, *quantity;
    uint32_t *extendedprice, *partkey, *suppkey, *revenue;
} mat_ctx;

static void mat_line(void *vctx, uint64_t row, const ssb_line *x) {
    mat_ctx *m = (mat_ctx *)vctx;
    if (m->orderdate) m->orderdate[row] = x->orderdate;
    if (m->discount) m->discount[row] = x->discount;
    if (m->quantity) m->quantity[row] = x->quantity;
    if (m->extendedprice) m->extendedprice[row] = x->extendedprice;
    if (m->partkey) m->partkey[row] = x->partkey;
    if (m->suppkey) m->suppkey[row] = x->suppkey;
    if (m->revenue) m->revenue[row] = x->revenue;
}

uint64_t oracle_ssb_generate(uint32_t sf, int32_t *lo_orderdate, uint8_t *lo_discount,
                             uint8_t *lo_quantity, uint32_t *lo_extendedprice,
                             uint32_t *lo_partkey, uint32_t *lo_suppkey, uint32_t *lo_revenue,
                             int32_t *d_key, int16_t *d_year, int32_t *d_yearmonthnum,
                             int16_t *d_weeknuminyear, int16_t *p_category, int16_t *p_brand1,
                             uint8_t *s_region) {
    ssb_state s;
    if (ssb_init(&s, sf)) return 0;
    mat_ctx m = {lo_orderdate, lo_discount, lo_quantity, lo_extendedprice,
                 lo_partkey,   lo_suppkey,  lo_revenue};
    uint64_t rows = gen_lineorder(&s, mat_line, &m);
    if (d_key) memcpy(d_key, s.dates.key, sizeof(s.dates.key));
    if (d_year) memcpy(d_year, s.dates.year, sizeof(s.dates.year));
    if (d_yearmonthnum) memcpy(d_yearmonthnum, s.dates.yearmonthnum, sizeof(s.dates.yearmonthnum));
    if (d_weeknuminyear)
        memcpy(d_weeknuminyear, s.dates.weeknuminyear, sizeof(s.dates.weeknuminyear));
    if (p_category) memcpy(p_category, s.p_category, s.parts * sizeof(int16_t));
    if (p_brand1) memcpy(p_brand1, s.p_brand1, s.parts * sizeof(int16_t));
    if (s_region) memcpy(s_region, s.s_region, s.suppliers);
    ssb_free(&s);
    return rows;
}
