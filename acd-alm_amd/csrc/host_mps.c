/* host_mps.c -- MPS input (acd_gpu.h: acd_mps_parse, acd_alm_read_mps).
 *
 * The reference reads LPs only in its own text format (alm.c:324-364);
 * MPS is the interchange format of LP test sets, and reading it is host
 * text parsing.  Fields are separated by white space (free MPS; fixed-
 * format files whose names contain no blanks read the same).  The LP is
 * turned into the ALM form min c'x s.t. Ax = b, l <= x <= u:
 *
 *   E row            a'x = b             (RANGES R: b <= a'x <= b + |R|
 *                                          if R > 0, b - |R| <= a'x <= b
 *                                          if R < 0)
 *   G row            a'x >= b            (b <= a'x <= b + |R|)
 *   L row            a'x <= b            (b - |R| <= a'x <= b)
 *
 * A row whose interval [lo, hi] is not a single point gets a slack column
 * after the structural ones: a'x - s = lo with 0 <= s <= hi - lo when lo
 * is finite, a'x + s = hi with s >= 0 otherwise.  Free rows (N rows after
 * the first, the objective) are dropped.  Column bounds default to
 * [0, inf); BOUNDS types UP, LO, FX, FR, MI, PL, BV, LI, UI (UP with a
 * negative value on a column whose lower bound is still the default
 * makes the lower bound -inf, the usual convention).  OBJSENSE MAX
 * negates the costs (and the objective constant) so the LP is still a
 * minimisation.
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acd_host.h"

/* ---- name table: open addressing, FNV-1a ---- */
struct names {
        char ** key;
        uint32_t * val;
        size_t cap, len;
};

static uint64_t hash_str(const char * s)
{
        uint64_t h = 1469598103934665603ull;
        for (; *s; s++) h = (h ^ (unsigned char)*s) * 1099511628211ull;
        return h;
}

static void names_free(struct names * t)
{
        for (size_t i = 0; i < t->cap; i++) free(t->key ? t->key[i] : NULL);
        free(t->key);
        free(t->val);
        memset(t, 0, sizeof *t);
}

static int names_grow(struct names * t)
{
        size_t nc = t->cap ? 2 * t->cap : 1024;
        char ** nk = calloc(nc, sizeof(char *));
        uint32_t * nv = calloc(nc, sizeof(uint32_t));
        if (nk == NULL || nv == NULL) {
                free(nk);
                free(nv);
                return -1;
        }
        for (size_t i = 0; i < t->cap; i++) {
                if (t->key[i] == NULL) continue;
                size_t j = hash_str(t->key[i]) & (nc - 1);
                while (nk[j] != NULL) j = (j + 1) & (nc - 1);
                nk[j] = t->key[i];
                nv[j] = t->val[i];
        }
        free(t->key);
        free(t->val);
        t->key = nk;
        t->val = nv;
        t->cap = nc;
        return 0;
}

/* Index of name, or -1. */
static int64_t names_find(const struct names * t, const char * name)
{
        if (t->cap == 0) return -1;
        size_t j = hash_str(name) & (t->cap - 1);
        while (t->key[j] != NULL) {
                if (strcmp(t->key[j], name) == 0) return t->val[j];
                j = (j + 1) & (t->cap - 1);
        }
        return -1;
}

/* Adds name with the next index; -1 on error or duplicate. */
static int64_t names_add(struct names * t, const char * name)
{
        if (names_find(t, name) >= 0) return -1;
        if (2 * (t->len + 1) > t->cap && names_grow(t)) return -1;
        size_t j = hash_str(name) & (t->cap - 1);
        while (t->key[j] != NULL) j = (j + 1) & (t->cap - 1);
        t->key[j] = strdup(name);
        if (t->key[j] == NULL) return -1;
        t->val[j] = (uint32_t)t->len;
        return (int64_t)t->len++;
}

/* ---- parser state ---- */
enum section { S_NONE, S_ROWS, S_COLUMNS, S_RHS, S_RANGES, S_BOUNDS,
               S_OBJSENSE, S_END };

struct parse {
        struct names rown, coln;
        char * rtype;           /* per row: 'E', 'L', 'G', 'N' */
        size_t nrow, caprow;
        double * rhs, * range;
        uint8_t * has_range;
        int64_t obj;            /* objective row, -1 until seen */
        double obj_rhs;
        /* triplets */
        uint32_t * tr, * tc;
        double * tv;
        size_t nnz, capnz;
        double * cost, * lo, * up;
        uint8_t * lo_set;
        size_t ncol, capcol;
        int maximize;
};

/* Appends one (row, column, value) entry, growing the three arrays. */
static int push_triplet(struct parse * p, uint32_t r, uint32_t c, double v)
{
        if (p->nnz == p->capnz) {
                size_t nc = p->capnz ? 2 * p->capnz : 1024;
                uint32_t * tr = realloc(p->tr, nc * sizeof(uint32_t));
                if (tr == NULL) return -1;
                p->tr = tr;
                uint32_t * tc = realloc(p->tc, nc * sizeof(uint32_t));
                if (tc == NULL) return -1;
                p->tc = tc;
                double * tv = realloc(p->tv, nc * sizeof(double));
                if (tv == NULL) return -1;
                p->tv = tv;
                p->capnz = nc;
        }
        p->tr[p->nnz] = r;
        p->tc[p->nnz] = c;
        p->tv[p->nnz] = v;
        p->nnz++;
        return 0;
}

static int split(char * line, char ** tok, int max)
{
        int n = 0;
        char * p = line;
        while (*p && n < max) {
                while (*p && isspace((unsigned char)*p)) p++;
                if (!*p) break;
                tok[n++] = p;
                while (*p && !isspace((unsigned char)*p)) p++;
                if (*p) *p++ = 0;
        }
        return n;
}

static int parse_double(const char * s, double * out)
{
        char * end;
        *out = strtod(s, &end);
        return end == s || *end != 0 ? -1 : 0;
}

static int ensure_row(struct parse * p, size_t n)
{
        if (n <= p->caprow) return 0;
        size_t nc = p->caprow ? p->caprow : 256;
        while (nc < n) nc *= 2;
        char * rt = realloc(p->rtype, nc);
        if (rt == NULL) return -1;
        p->rtype = rt;
        double * rh = realloc(p->rhs, nc * sizeof(double));
        if (rh == NULL) return -1;
        p->rhs = rh;
        double * rg = realloc(p->range, nc * sizeof(double));
        if (rg == NULL) return -1;
        p->range = rg;
        uint8_t * hr = realloc(p->has_range, nc);
        if (hr == NULL) return -1;
        p->has_range = hr;
        p->caprow = nc;
        return 0;
}

static int ensure_col(struct parse * p, size_t n)
{
        if (n <= p->capcol) return 0;
        size_t nc = p->capcol ? p->capcol : 256;
        while (nc < n) nc *= 2;
        double * c = realloc(p->cost, nc * sizeof(double));
        if (c == NULL) return -1;
        p->cost = c;
        double * l = realloc(p->lo, nc * sizeof(double));
        if (l == NULL) return -1;
        p->lo = l;
        double * u = realloc(p->up, nc * sizeof(double));
        if (u == NULL) return -1;
        p->up = u;
        uint8_t * s = realloc(p->lo_set, nc);
        if (s == NULL) return -1;
        p->lo_set = s;
        p->capcol = nc;
        return 0;
}

static void parse_free(struct parse * p)
{
        names_free(&p->rown);
        names_free(&p->coln);
        free(p->rtype);
        free(p->rhs);
        free(p->range);
        free(p->has_range);
        free(p->tr);
        free(p->tc);
        free(p->tv);
        free(p->cost);
        free(p->lo);
        free(p->up);
        free(p->lo_set);
        memset(p, 0, sizeof *p);
}

#define FAIL(...)                                                        \
        do {                                                             \
                snprintf(out->error, sizeof out->error, __VA_ARGS__);    \
                goto fail;                                               \
        } while (0)

/* Pairs (row name, value) of a RHS / RANGES line after the set name. */
static int row_pairs(struct parse * p, char ** tok, int n, int ranges,
                     struct acd_mps_lp * out, size_t lineno)
{
        for (int k = 0; k + 1 < n; k += 2) {
                double v;
                if (parse_double(tok[k + 1], &v)) {
                        snprintf(out->error, sizeof out->error,
                                 "line %zu: bad number '%s'", lineno, tok[k + 1]);
                        return -1;
                }
                int64_t r = names_find(&p->rown, tok[k]);
                if (r < 0) {
                        snprintf(out->error, sizeof out->error,
                                 "line %zu: unknown row '%s'", lineno, tok[k]);
                        return -1;
                }
                if (r == p->obj) {
                        if (!ranges) p->obj_rhs = v;
                        continue;
                }
                if (ranges) {
                        p->range[r] = v;
                        p->has_range[r] = 1;
                } else {
                        p->rhs[r] = v;
                }
        }
        return 0;
}

/* Interval [lo, hi] of a kept row: E is b (a RANGES value R widens it to
 * [b, b + |R|] for R > 0, [b - |R|, b] for R < 0), G is [b, b + |R|] or
 * [b, inf), L is [b - |R|, b] or (-inf, b].  A G or L row with RANGES 0 is
 * the point b (no slack). */
static void row_interval(char type, double b, int has_range, double range,
                         double * lo, double * hi)
{
        const double R = fabs(range);
        switch (type) {
        case 'E':
                *lo = *hi = b;
                if (has_range && range > 0) *hi = b + R;
                if (has_range && range < 0) *lo = b - R;
                break;
        case 'G':
                *lo = b;
                *hi = has_range ? b + R : HUGE_VAL;
                break;
        default:  /* 'L' */
                *lo = has_range ? b - R : -HUGE_VAL;
                *hi = b;
                break;
        }
}

int acd_mps_parse(FILE * stream, struct acd_mps_lp * out)
{
        struct parse p;
        memset(&p, 0, sizeof p);
        memset(out, 0, sizeof *out);
        p.obj = -1;
        enum section sec = S_NONE;
        char * line = NULL;
        size_t linecap = 0, lineno = 0;
        ssize_t len;
        int64_t * rowmap = NULL;
        uint32_t * rows = NULL, * cols = NULL;
        double * vals = NULL, * rhs = NULL, * lin = NULL, * lo = NULL,
               * up = NULL;
        if (stream == NULL) {
                snprintf(out->error, sizeof out->error, "no stream");
                return -1;
        }
        while ((len = getline(&line, &linecap, stream)) >= 0) {
                lineno++;
                if (line[0] == '*' || line[0] == '\n' || line[0] == '\r'
                    || line[0] == 0)
                        continue;
                char * tok[8];
                const int header = !isspace((unsigned char)line[0]);
                int n = split(line, tok, 8);
                if (n == 0) continue;
                if (header) {
                        if (!strcmp(tok[0], "NAME")) sec = S_NONE;
                        else if (!strcmp(tok[0], "ROWS")) sec = S_ROWS;
                        else if (!strcmp(tok[0], "COLUMNS")) sec = S_COLUMNS;
                        else if (!strcmp(tok[0], "RHS")) sec = S_RHS;
                        else if (!strcmp(tok[0], "RANGES")) sec = S_RANGES;
                        else if (!strcmp(tok[0], "BOUNDS")) sec = S_BOUNDS;
                        else if (!strcmp(tok[0], "ENDATA")) {
                                sec = S_END;
                                break;
                        } else if (!strcmp(tok[0], "OBJSENSE")) {
                                /* the sense on the same line or the next */
                                sec = S_OBJSENSE;
                                if (n > 1) {
                                        p.maximize = !strcmp(tok[1], "MAX")
                                                     || !strcmp(tok[1], "MAXIMIZE");
                                        sec = S_NONE;
                                }
                        } else {
                                FAIL("line %zu: unknown section '%s'", lineno,
                                     tok[0]);
                        }
                        continue;
                }
                switch (sec) {
                case S_OBJSENSE:
                        p.maximize = !strcmp(tok[0], "MAX")
                                     || !strcmp(tok[0], "MAXIMIZE");
                        sec = S_NONE;
                        break;
                case S_ROWS: {
                        if (n < 2 || strlen(tok[0]) != 1
                            || !strchr("NELG", tok[0][0]))
                                FAIL("line %zu: bad ROWS entry", lineno);
                        int64_t r = names_add(&p.rown, tok[1]);
                        if (r < 0) FAIL("line %zu: duplicate row '%s'", lineno, tok[1]);
                        if (ensure_row(&p, (size_t)r + 1))
                                FAIL("out of memory");
                        p.rtype[r] = tok[0][0];
                        p.rhs[r] = 0;
                        p.range[r] = 0;
                        p.has_range[r] = 0;
                        p.nrow = (size_t)r + 1;
                        if (tok[0][0] == 'N' && p.obj < 0) p.obj = r;
                        break;
                }
                case S_COLUMNS: {
                        if (n >= 3 && strstr(tok[1], "MARKER") != NULL)
                                break;  /* integrality markers: ignored */
                        if (n < 3 || n % 2 == 0)
                                FAIL("line %zu: bad COLUMNS entry", lineno);
                        int64_t c = names_find(&p.coln, tok[0]);
                        if (c < 0) {
                                c = names_add(&p.coln, tok[0]);
                                if (c < 0 || ensure_col(&p, (size_t)c + 1))
                                        FAIL("out of memory");
                                p.cost[c] = 0;
                                p.lo[c] = 0;
                                p.up[c] = HUGE_VAL;
                                p.lo_set[c] = 0;
                                p.ncol = (size_t)c + 1;
                        }
                        for (int k = 1; k + 1 < n; k += 2) {
                                double v;
                                if (parse_double(tok[k + 1], &v))
                                        FAIL("line %zu: bad number '%s'", lineno,
                                             tok[k + 1]);
                                int64_t r = names_find(&p.rown, tok[k]);
                                if (r < 0)
                                        FAIL("line %zu: unknown row '%s'", lineno,
                                             tok[k]);
                                if (r == p.obj) {
                                        p.cost[c] += v;
                                        continue;
                                }
                                if (p.rtype[r] == 'N' || v == 0) continue;
                                if (push_triplet(&p, (uint32_t)r, (uint32_t)c, v))
                                        goto oom;
                        }
                        break;
                }
                case S_RHS:
                case S_RANGES: {
                        /* optional set name: an odd token count has one */
                        const int skip = n % 2;
                        if (n - skip < 2)
                                FAIL("line %zu: bad %s entry", lineno,
                                     sec == S_RHS ? "RHS" : "RANGES");
                        if (row_pairs(&p, tok + skip, n - skip, sec == S_RANGES,
                                      out, lineno))
                                goto fail;
                        break;
                }
                case S_BOUNDS: {
                        const char * t = tok[0];
                        const int novalue = !strcmp(t, "FR") || !strcmp(t, "MI")
                                            || !strcmp(t, "PL") || !strcmp(t, "BV");
                        /* type [set] column [value] */
                        const int want = novalue ? 2 : 3;
                        int ci;
                        if (n == want) ci = 1;
                        else if (n == want + 1) ci = 2;
                        else if (!strcmp(t, "BV") && n == 4) ci = 2;
                        else FAIL("line %zu: bad BOUNDS entry", lineno);
                        int64_t c = names_find(&p.coln, tok[ci]);
                        if (c < 0) FAIL("line %zu: unknown column '%s'", lineno, tok[ci]);
                        double v = 0;
                        if (!novalue && parse_double(tok[ci + 1], &v))
                                FAIL("line %zu: bad number '%s'", lineno, tok[ci + 1]);
                        if (!strcmp(t, "UP") || !strcmp(t, "UI")) {
                                p.up[c] = v;
                                if (v < 0 && !p.lo_set[c]) p.lo[c] = -HUGE_VAL;
                        } else if (!strcmp(t, "LO") || !strcmp(t, "LI")) {
                                p.lo[c] = v;
                                p.lo_set[c] = 1;
                        } else if (!strcmp(t, "FX")) {
                                p.lo[c] = p.up[c] = v;
                                p.lo_set[c] = 1;
                        } else if (!strcmp(t, "FR")) {
                                p.lo[c] = -HUGE_VAL;
                                p.up[c] = HUGE_VAL;
                                p.lo_set[c] = 1;
                        } else if (!strcmp(t, "MI")) {
                                p.lo[c] = -HUGE_VAL;
                                p.lo_set[c] = 1;
                        } else if (!strcmp(t, "PL")) {
                                p.up[c] = HUGE_VAL;
                        } else if (!strcmp(t, "BV")) {
                                p.lo[c] = 0;
                                p.up[c] = 1;
                                p.lo_set[c] = 1;
                        } else {
                                FAIL("line %zu: unsupported bound type '%s'",
                                     lineno, t);
                        }
                        break;
                }
                default:
                        break;  /* NAME line contents and the like */
                }
        }
        if (sec != S_END) FAIL("missing ENDATA");
        if (p.obj < 0) FAIL("no objective (N) row");

        /* rows kept (E/L/G), their intervals, slacks */
        rowmap = malloc((p.nrow + 1) * sizeof(int64_t));
        if (rowmap == NULL) goto oom;
        size_t m = 0, nslack = 0;
        for (size_t r = 0; r < p.nrow; r++) {
                rowmap[r] = -1;
                if (p.rtype[r] == 'N') continue;
                rowmap[r] = (int64_t)m++;
                double lo_r, hi_r;
                row_interval(p.rtype[r], p.rhs[r], p.has_range[r],
                             p.range[r], &lo_r, &hi_r);
                if (lo_r != hi_r) nslack++;   /* the test of the fill below */
        }
        const size_t n = p.ncol + nslack;
        const size_t nnz = p.nnz + nslack;
        rows = malloc((nnz + 1) * sizeof(uint32_t));
        cols = malloc((nnz + 1) * sizeof(uint32_t));
        vals = malloc((nnz + 1) * sizeof(double));
        rhs = malloc((m + 1) * sizeof(double));
        lin = malloc((n + 1) * sizeof(double));
        lo = malloc((n + 1) * sizeof(double));
        up = malloc((n + 1) * sizeof(double));
        if (!rows || !cols || !vals || !rhs || !lin || !lo || !up) goto oom;
        if (m > UINT32_MAX || n > UINT32_MAX) FAIL("LP too large");
        for (size_t k = 0; k < p.nnz; k++) {
                rows[k] = (uint32_t)rowmap[p.tr[k]];
                cols[k] = p.tc[k];
                vals[k] = p.tv[k];
        }
        const double sign = p.maximize ? -1.0 : 1.0;
        for (size_t c = 0; c < p.ncol; c++) {
                lin[c] = sign * p.cost[c];
                lo[c] = p.lo[c];
                up[c] = p.up[c];
        }
        size_t s = p.ncol, k = p.nnz;
        for (size_t r = 0; r < p.nrow; r++) {
                if (rowmap[r] < 0) continue;
                const size_t i = (size_t)rowmap[r];
                double lo_r, hi_r;
                row_interval(p.rtype[r], p.rhs[r], p.has_range[r],
                             p.range[r], &lo_r, &hi_r);
                if (lo_r == hi_r) {
                        rhs[i] = lo_r;
                        continue;
                }
                /* slack column s */
                rows[k] = (uint32_t)i;
                cols[k] = (uint32_t)s;
                lin[s] = 0;
                lo[s] = 0;
                if (isfinite(lo_r)) {   /* a'x - s = lo, 0 <= s <= hi - lo */
                        vals[k] = -1.0;
                        rhs[i] = lo_r;
                        up[s] = isfinite(hi_r) ? hi_r - lo_r : HUGE_VAL;
                } else {                /* a'x + s = hi, s >= 0 */
                        vals[k] = 1.0;
                        rhs[i] = hi_r;
                        up[s] = HUGE_VAL;
                }
                k++;
                s++;
        }
        if (s != n || k != nnz) FAIL("internal: slack count mismatch");
        out->nrows = m;
        out->ncols = n;
        out->nnz = nnz;
        out->nstructural = p.ncol;
        out->rows = rows;
        out->cols = cols;
        out->values = vals;
        out->rhs = rhs;
        out->linear = lin;
        out->lower = lo;
        out->upper = up;
        out->objective_offset = -sign * p.obj_rhs;
        out->maximize = p.maximize;
        free(rowmap);
        free(line);
        parse_free(&p);
        return 0;
oom:
        snprintf(out->error, sizeof out->error, "out of memory");
fail:
        free(rowmap);
        free(rows);
        free(cols);
        free(vals);
        free(rhs);
        free(lin);
        free(lo);
        free(up);
        free(line);
        parse_free(&p);
        return -1;
}

void acd_mps_free(struct acd_mps_lp * lp)
{
        if (lp == NULL) return;
        free(lp->rows);
        free(lp->cols);
        free(lp->values);
        free(lp->rhs);
        free(lp->linear);
        free(lp->lower);
        free(lp->upper);
        lp->rows = lp->cols = NULL;
        lp->values = lp->rhs = lp->linear = lp->lower = lp->upper = NULL;
}

alm_t * acd_alm_read_mps(FILE * stream, struct acd_mps_lp * OUT_info)
{
        struct acd_mps_lp lp;
        if (acd_mps_parse(stream, &lp)) {
                if (OUT_info != NULL) *OUT_info = lp;
                return NULL;
        }
        alm_t * alm = NULL;
        sparse_matrix_t * m = sparse_matrix_make(lp.ncols, lp.nrows, lp.nnz,
                                                 lp.rows, lp.cols, lp.values, 0);
        if (m != NULL) {
                /* every row is an equality: free multipliers */
                double * llo = malloc((lp.nrows + 1) * sizeof(double));
                double * lhi = malloc((lp.nrows + 1) * sizeof(double));
                if (llo != NULL && lhi != NULL) {
                        for (size_t i = 0; i < lp.nrows; i++) {
                                llo[i] = -HUGE_VAL;
                                lhi[i] = HUGE_VAL;
                        }
                        alm = alm_make(m, lp.nrows, lp.rhs, lp.ncols, lp.linear,
                                       lp.lower, lp.upper, llo, lhi);
                }
                free(llo);
                free(lhi);
                if (alm == NULL) sparse_matrix_free(m);
        }
        if (alm == NULL && lp.error[0] == 0)
                snprintf(lp.error, sizeof lp.error, "matrix or ALM set-up failed");
        acd_mps_free(&lp);
        if (OUT_info != NULL) *OUT_info = lp;
        return alm;
}
