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

/* Nation, city and region as integers: nation = index into the sorted nation names, city =
 * nation * 10 + city_code (the city string is the nation name's first nine characters followed by
 * the code, SSB_def.hpp:335,362, and those prefixes differ between nations, so the mapping is
 * one to one), region = index into AFRICA, AMERICA, ASIA, EUROPE, MIDDLE EAST. */
typedef struct {
    uint64_t seed;
    uint64_t orders, parts, suppliers, customers;
    ssb_dates dates;
    int16_t *p_category, *p_brand1;
    uint8_t *p_mfgr;
    uint8_t *s_region, *s_nation;
    int16_t *s_city;
    uint8_t *c_region, *c_nation;
    int16_t *c_city;
} ssb_state;

/* populate_customer (SSB_def.hpp:344-370): 6 draws per row. */
static void gen_customers(ssb_state *s) {
    for (uint64_t i = 0; i < s->customers; ++i) {
        uint32_t nation = lcg_uniform(&s->seed, 0, 25);
        uint32_t city = lcg_uniform(&s->seed, 0, 10);
        lcg_uniform(&s->seed, 100, 1000); /* acode */
        lcg_uniform(&s->seed, 100, 1000); /* exchg */
        lcg_uniform(&s->seed, 1000, 10000); /* pnumber */
        lcg_uniform(&s->seed, 0, N_SEGMENT); /* mktsegment */
        s->c_nation[i] = (uint8_t)nation;
        s->c_city[i] = (int16_t)(nation * 10 + city);
        s->c_region[i] = (uint8_t)NATION_REGION_SORTED[nation];
    }
}

/* populate_supplier (SSB_def.hpp:318-342): 5 draws per row. */
static void gen_suppliers(ssb_state *s) {
    for (uint64_t i = 0; i < s->suppliers; ++i) {
        uint32_t nation = lcg_uniform(&s->seed, 0, 25);
        uint32_t city = lcg_uniform(&s->seed, 0, 10);
        lcg_uniform(&s->seed, 100, 1000);
        lcg_uniform(&s->seed, 100, 1000);
        lcg_uniform(&s->seed, 1000, 10000);
        s->s_region[i] = (uint8_t)NATION_REGION_SORTED[nation];
        s->s_nation[i] = (uint8_t)nation;
        s->s_city[i] = (int16_t)(nation * 10 + city);
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
        s->p_mfgr[i] = (uint8_t)num1; /* "MFGR#<num1>", SSB_def.hpp:302 */
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
    s->p_mfgr = (uint8_t *)malloc(s->parts);
    s->s_region = (uint8_t *)malloc(s->suppliers);
    s->s_nation = (uint8_t *)malloc(s->suppliers);
    s->s_city = (int16_t *)malloc(s->suppliers * sizeof(int16_t));
    s->c_region = (uint8_t *)malloc(s->customers);
    s->c_nation = (uint8_t *)malloc(s->customers);
    s->c_city = (int16_t *)malloc(s->customers * sizeof(int16_t));
    if (!s->p_category || !s->p_brand1 || !s->p_mfgr || !s->s_region || !s->s_nation || !s->s_city ||
        !s->c_region || !s->c_nation || !s->c_city)
        return -1;
    gen_customers(s);
    gen_suppliers(s);
    gen_parts(s);
    return 0;
}

static void ssb_free(ssb_state *s) {
    free(s->p_category);
    free(s->p_brand1);
    free(s->p_mfgr);
    free(s->s_region);
    free(s->s_nation);
    free(s->s_city);
    free(s->c_region);
    free(s->c_nation);
    free(s->c_city);
}

typedef struct {
    int32_t orderdate;
    uint32_t date_idx;
    uint8_t quantity, discount;
    uint32_t extendedprice, revenue, supplycost;
    uint32_t partkey, suppkey, custkey;
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
            x.supplycost = 6 * partkey_to_price(x.partkey) / 10; /* SSB_def.hpp:252 */
            x.custkey = cust;
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
/* Q3.x / Q4.x group tables (dense, indexed by the integer group attributes) */
#define Q3N (25u * 25u * Q2_YEARS)   /* Q3.1: (c_nation, s_nation, d_year) */
#define Q3C (250u * 250u * Q2_YEARS) /* Q3.2..Q3.4: (c_city, s_city, d_year) */
#define Q41 (Q2_YEARS * 25u)         /* (d_year, c_nation) */
#define Q42 (Q2_YEARS * 25u * 56u)   /* (d_year, s_nation, p_category) */
#define Q43 (Q2_YEARS * 250u * 41u)  /* (d_year, s_city, p_brand1 - 1400) for category MFGR#14 */
static const uint32_t Q34_SLOTS[7] = {Q3N, Q3C, Q3C, Q3C, Q41, Q42, Q43};
typedef struct {
    ssb_state *s;
    uint64_t q1[3];
    uint64_t *q2sum[3];
    uint64_t *q2cnt[3];
    uint64_t *qsum[7]; /* Q3.1..Q3.4, Q4.1..Q4.3 */
    uint64_t *qcnt[7];
} golden_ctx;

static inline void q34_add(golden_ctx *g, int q, uint32_t slot, uint64_t v) {
    g->qsum[q][slot] += v;
    g->qcnt[q][slot] += 1;
}

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

    /* Q3.x (03.hpp:611-677, pipeline_q3 03.cuh:118-170): customer, supplier and date
     * predicates, SUM(lo_revenue) by (customer attribute, supplier attribute, d_year) */
    const ssb_state *s = g->s;
    const uint32_t c = x->custkey, sp = x->suppkey, y = (uint32_t)(d->year[di] - 1992);
    const int y9297 = d->year[di] >= 1992 && d->year[di] <= 1997;
    const int cki = s->c_city[c] == 221 || s->c_city[c] == 225; /* UNITED KI1, UNITED KI5 */
    const int ski = s->s_city[sp] == 221 || s->s_city[sp] == 225;
    if (s->c_region[c] == 2 && s->s_region[sp] == 2 && y9297) /* ASIA, ASIA */
        q34_add(g, 0, ((uint32_t)s->c_nation[c] * 25u + s->s_nation[sp]) * Q2_YEARS + y, x->revenue);
    const uint32_t cs = ((uint32_t)s->c_city[c] * 250u + (uint32_t)s->s_city[sp]) * Q2_YEARS + y;
    if (s->c_nation[c] == 23 && s->s_nation[sp] == 23 && y9297) /* UNITED STATES */
        q34_add(g, 1, cs, x->revenue);
    if (cki && ski && y9297) q34_add(g, 2, cs, x->revenue);
    if (cki && ski && d->yearmonthnum[di] == 199712) q34_add(g, 3, cs, x->revenue); /* Dec1997 */

    /* Q4.x (04.hpp:316-381, pipeline_q4 04.cuh:99-153): SUM(lo_revenue - lo_supplycost), the
     * difference taken in 32-bit unsigned arithmetic as the reference's uint32 columns give it */
    const uint64_t profit = (uint32_t)(x->revenue - x->supplycost);
    const uint32_t mf = s->p_mfgr[x->partkey];
    const int y9798 = d->year[di] == 1997 || d->year[di] == 1998;
    if (s->c_region[c] == 1 && s->s_region[sp] == 1 && (mf == 1 || mf == 2)) { /* AMERICA, MFGR#1/2 */
        q34_add(g, 4, y * 25u + s->c_nation[c], profit);
        if (y9798) q34_add(g, 5, (y * 25u + s->s_nation[sp]) * 56u + (uint32_t)cat, profit);
    }
    if (s->c_region[c] == 1 && s->s_nation[sp] == 23 && y9798 && cat == 14) /* MFGR#14 */
        q34_add(g, 6, (y * 250u + (uint32_t)s->s_city[sp]) * 41u + (uint32_t)(brand - 1400), profit);
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
    for (int q = 0; q < 7; ++q) {
        g.qsum[q] = (uint64_t *)calloc(Q34_SLOTS[q], sizeof(uint64_t));
        g.qcnt[q] = (uint64_t *)calloc(Q34_SLOTS[q], sizeof(uint64_t));
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
    for (int q = 0; q < 7; ++q) {
        uint64_t sum = 0, groups = 0;
        for (uint32_t i = 0; i < Q34_SLOTS[q]; ++i) {
            if (g.qcnt[q][i]) {
                groups++;
                sum += g.qsum[q][i];
            }
        }
        out->q34_sum[q] = sum;
        out->q34_groups[q] = groups;
        free(g.qsum[q]);
        free(g.qcnt[q]);
    }
    ssb_free(&s);
    return 0;
}

static void mat_line(void *vctx, uint64_t row, const ssb_line *x) {
    const oracle_ssb_cols *m = (const oracle_ssb_cols *)vctx;
    if (m->lo_orderdate) m->lo_orderdate[row] = x->orderdate;
    if (m->lo_discount) m->lo_discount[row] = x->discount;
    if (m->lo_quantity) m->lo_quantity[row] = x->quantity;
    if (m->lo_extendedprice) m->lo_extendedprice[row] = x->extendedprice;
    if (m->lo_partkey) m->lo_partkey[row] = x->partkey;
    if (m->lo_suppkey) m->lo_suppkey[row] = x->suppkey;
    if (m->lo_custkey) m->lo_custkey[row] = x->custkey;
    if (m->lo_revenue) m->lo_revenue[row] = x->revenue;
    if (m->lo_supplycost) m->lo_supplycost[row] = x->supplycost;
}

#define COPY_OUT(dst, src, n) \
    do {                      \
        if (dst) memcpy(dst, src, n); \
    } while (0)
uint64_t oracle_ssb_generate(uint32_t sf, const oracle_ssb_cols *o) {
    ssb_state s;
    if (ssb_init(&s, sf)) return 0;
    uint64_t rows = gen_lineorder(&s, mat_line, (void *)o);
    COPY_OUT(o->d_key, s.dates.key, sizeof(s.dates.key));
    COPY_OUT(o->d_year, s.dates.year, sizeof(s.dates.year));
    COPY_OUT(o->d_yearmonthnum, s.dates.yearmonthnum, sizeof(s.dates.yearmonthnum));
    COPY_OUT(o->d_weeknuminyear, s.dates.weeknuminyear, sizeof(s.dates.weeknuminyear));
    COPY_OUT(o->p_category, s.p_category, s.parts * sizeof(int16_t));
    COPY_OUT(o->p_brand1, s.p_brand1, s.parts * sizeof(int16_t));
    COPY_OUT(o->p_mfgr, s.p_mfgr, s.parts);
    COPY_OUT(o->s_region, s.s_region, s.suppliers);
    COPY_OUT(o->s_nation, s.s_nation, s.suppliers);
    COPY_OUT(o->s_city, s.s_city, s.suppliers * sizeof(int16_t));
    COPY_OUT(o->c_region, s.c_region, s.customers);
    COPY_OUT(o->c_nation, s.c_nation, s.customers);
    COPY_OUT(o->c_city, s.c_city, s.customers * sizeof(int16_t));
    ssb_free(&s);
    return rows;
}
