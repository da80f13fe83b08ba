/*
 * sw2_case.c -- SW2 case-file front end.
 *
 * SW2 pre-processes <bname>.sw with the external m4 (sppde/sppde_parser.c:
 * 177-210), then reads mandatory tags in the fixed order of init.c:48-229.
 * Every "<expr>" on a line is evaluated (tinyexpr grammar, sppde_tinyexpr.c)
 * and re-printed with %.15e before the line is scanned
 * (sppde_parser.c:84-143).  This file does the same without external tools:
 *   - an m4 subset: define(NAME,TEXT), include(FILE), dnl, '#' comments
 *     copied without expansion, rescanning of expanded text;
 *   - a recursive-descent evaluator of the tinyexpr grammar ('^' left
 *     associative, unary minus binding tighter than '^', log = log10);
 *   - the tag reader (p_getint/p_getdouble/p_getndoubles/p_alloc_gettable,
 *     sppde_parser.c:236-317).
 */
#include <ctype.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sw2_internal.h"

/* ------------------------------------------------------------------------
 * expression evaluator (grammar of sppde_tinyexpr.c:380-470)
 *   list   = expr {"," expr}
 *   expr   = term {("+"|"-") term}
 *   term   = factor {("*"|"/"|"%") factor}
 *   factor = power {"^" power}
 *   power  = {("-"|"+")} base
 *   base   = number | fun0 ["(" ")"] | fun1 power | fun2 "(" expr "," expr ")"
 *          | "(" list ")"
 * ---------------------------------------------------------------------- */
typedef struct {
    const char *p;
    int err;
} xstate;

/* fac(a): the factorial of a's integer part, as tinyexpr's fac gives it:
 * NaN below zero, infinity once the value no longer fits in 64-bit unsigned
 * integers (from 21!).  The running product 2*3*...*k is exact in double up
 * to 20!: with its factors of two set aside, 20! has 44 significant bits. */
static double x_fac(double a) {
    if (a < 0.0) return NAN;
    if (!(a < 21.0)) return INFINITY;
    const int n = (int)a;
    double f = 1.0;
    for (int k = 2; k <= n; k++) f *= (double)k;
    return f;
}

/* ncr(n, r): the binomial coefficient of the integer parts, as tinyexpr's
 * ncr: NaN for a negative argument or n < r, infinity for arguments beyond
 * 32 bits or a value beyond 64 bits.  C(m, k) with k = min(r, m - r) is
 * built up along the diagonal C(m-k+i, i) = C(m-k+i-1, i-1) * (m-k+i) / i,
 * each step an exact division in 128-bit arithmetic. */
static double x_ncr(double n, double r) {
    if (n < 0.0 || r < 0.0 || n < r) return NAN;
    if (n > 4294967295.0 || r > 4294967295.0) return INFINITY;
    const uint64_t m = (uint64_t)n;
    uint64_t k = (uint64_t)r;
    if (k > m - k) k = m - k;
    unsigned __int128 c = 1;
    for (uint64_t i = 1; i <= k; i++) {
        c = c * (m - k + i) / i;
        if (c > (unsigned __int128)UINT64_MAX) return INFINITY;
    }
    return (double)(uint64_t)c;
}
static double x_pi(void) { return 3.14159265358979323846; }
static double x_e(void) { return 2.71828182845904523536; }

typedef struct {
    const char *name;
    int arity;
    double (*f0)(void);
    double (*f1)(double);
    double (*f2)(double, double);
} xfun;

static double x_npr(double n, double r) { return x_ncr(n, r) * x_fac(r); }

static const xfun XFUN[] = {
    {"abs", 1, 0, fabs, 0},     {"acos", 1, 0, acos, 0},   {"asin", 1, 0, asin, 0},   {"atan", 1, 0, atan, 0},
    {"atan2", 2, 0, 0, atan2},  {"ceil", 1, 0, ceil, 0},   {"cos", 1, 0, cos, 0},     {"cosh", 1, 0, cosh, 0},
    {"e", 0, x_e, 0, 0},        {"exp", 1, 0, exp, 0},     {"fac", 1, 0, x_fac, 0},   {"floor", 1, 0, floor, 0},
    {"ln", 1, 0, log, 0},       {"log", 1, 0, log10, 0},   {"log10", 1, 0, log10, 0}, {"ncr", 2, 0, 0, x_ncr},
    {"npr", 2, 0, 0, x_npr},    {"pi", 0, x_pi, 0, 0},     {"pow", 2, 0, 0, pow},     {"sin", 1, 0, sin, 0},
    {"sinh", 1, 0, sinh, 0},    {"sqrt", 1, 0, sqrt, 0},   {"tan", 1, 0, tan, 0},     {"tanh", 1, 0, tanh, 0},
};

static void x_ws(xstate *s) {
    while (*s->p == ' ' || *s->p == '\t' || *s->p == '\n' || *s->p == '\r') s->p++;
}
static double x_list(xstate *s);
static double x_expr(xstate *s);
static double x_power(xstate *s);

static double x_base(xstate *s) {
    x_ws(s);
    const char *q = s->p;
    if ((*q >= '0' && *q <= '9') || *q == '.') {
        char *end;
        double v = strtod(q, &end);
        s->p = end;
        return v;
    }
    if (isalpha((unsigned char)*q)) {
        const char *st = q;
        while (isalpha((unsigned char)*q) || isdigit((unsigned char)*q) || *q == '_') q++;
        size_t len = (size_t)(q - st);
        const xfun *fn = NULL;
        for (size_t k = 0; k < sizeof XFUN / sizeof XFUN[0]; k++)
            if (strlen(XFUN[k].name) == len && strncmp(XFUN[k].name, st, len) == 0) fn = &XFUN[k];
        s->p = q;
        if (!fn) {
            s->err = 1;
            return NAN;
        }
        if (fn->arity == 0) {
            x_ws(s);
            if (*s->p == '(') {
                s->p++;
                x_ws(s);
                if (*s->p != ')') { s->err = 1; return NAN; }
                s->p++;
            }
            return fn->f0();
        }
        if (fn->arity == 1) return fn->f1(x_power(s));
        x_ws(s);
        if (*s->p != '(') { s->err = 1; return NAN; }
        s->p++;
        double a = x_expr(s);
        x_ws(s);
        if (*s->p != ',') { s->err = 1; return NAN; }
        s->p++;
        double b = x_expr(s);
        x_ws(s);
        if (*s->p != ')') { s->err = 1; return NAN; }
        s->p++;
        return fn->f2(a, b);
    }
    if (*q == '(') {
        s->p++;
        double v = x_list(s);
        x_ws(s);
        if (*s->p != ')') { s->err = 1; return NAN; }
        s->p++;
        return v;
    }
    s->err = 1;
    return NAN;
}

static double x_power(xstate *s) {
    int sign = 1;
    for (;;) {
        x_ws(s);
        if (*s->p == '-') { sign = -sign; s->p++; }
        else if (*s->p == '+') s->p++;
        else break;
    }
    double b = x_base(s);
    return sign == 1 ? b : -b;
}

static double x_factor(xstate *s) {
    double v = x_power(s);
    for (;;) {
        x_ws(s);
        if (*s->p != '^') break;
        s->p++;
        v = pow(v, x_power(s));
    }
    return v;
}

static double x_term(xstate *s) {
    double v = x_factor(s);
    for (;;) {
        x_ws(s);
        char c = *s->p;
        if (c != '*' && c != '/' && c != '%') break;
        s->p++;
        double r = x_factor(s);
        if (c == '*') v = v * r;
        else if (c == '/') v = v / r;
        else v = fmod(v, r);
    }
    return v;
}

static double x_expr(xstate *s) {
    double v = x_term(s);
    for (;;) {
        x_ws(s);
        char c = *s->p;
        if (c != '+' && c != '-') break;
        s->p++;
        double r = x_term(s);
        v = (c == '+') ? v + r : v - r;
    }
    return v;
}

static double x_list(xstate *s) {
    double v = x_expr(s);
    for (;;) {
        x_ws(s);
        if (*s->p != ',') break;
        s->p++;
        v = x_expr(s);
    }
    return v;
}

double sw2_eval_expr(const char *expr, int *error) {
    xstate s = {expr, 0};
    double v = x_list(&s);
    x_ws(&s);
    if (*s.p != '\0') s.err = 1;
    if (error) *error = s.err;
    return s.err ? NAN : v;
}

/* ------------------------------------------------------------------------
 * m4 subset
 * ---------------------------------------------------------------------- */
typedef struct {
    char name[128];
    char *text;
} m4def;

typedef struct {
    m4def defs[256];
    int ndefs;
    char *out;
    size_t len, cap;
    char dir[1024]; /* directory of the top case file, second include path */
    char err[512];
} m4state;

static void m4_put(m4state *m, const char *s, size_t n) {
    if (m->len + n + 1 > m->cap) {
        m->cap = (m->len + n + 1) * 2 + 4096;
        m->out = (char *)realloc(m->out, m->cap);
    }
    memcpy(m->out + m->len, s, n);
    m->len += n;
    m->out[m->len] = '\0';
}

static char *slurp(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *b = (char *)malloc((size_t)n + 1);
    size_t got = fread(b, 1, (size_t)n, f);
    b[got] = '\0';
    fclose(f);
    return b;
}

static int m4_run(m4state *m, const char *txt, int depth);

/* collect the arguments of NAME( ... ) honouring nested parentheses;
 * returns pointer after ')' or NULL */
static const char *m4_args(const char *p, char args[][4096], int *nargs) {
    int lvl = 0, k = 0;
    size_t pos = 0;
    *nargs = 0;
    p++; /* skip '(' */
    while (*p == ' ' || *p == '\t' || *p == '\n') p++; /* m4 drops leading blanks */
    for (; *p; p++) {
        if (*p == '(') lvl++;
        if (*p == ')') {
            if (lvl == 0) {
                args[k][pos] = '\0';
                *nargs = k + 1;
                return p + 1;
            }
            lvl--;
        }
        if (*p == ',' && lvl == 0) {
            args[k][pos] = '\0';
            k++;
            pos = 0;
            if (k >= 9) return NULL;
            while (p[1] == ' ' || p[1] == '\t' || p[1] == '\n') p++;
            continue;
        }
        if (pos < 4095) args[k][pos++] = *p;
    }
    return NULL;
}

static int m4_run(m4state *m, const char *txt, int depth) {
    if (depth > 64) {
        snprintf(m->err, sizeof m->err, "m4: recursion too deep");
        return -1;
    }
    const char *p = txt;
    while (*p) {
        if (*p == '#') { /* comment: copied verbatim, never expanded */
            const char *e = strchr(p, '\n');
            size_t n = e ? (size_t)(e - p) : strlen(p);
            m4_put(m, p, n);
            p += n;
            continue;
        }
        if (isalpha((unsigned char)*p) || *p == '_') {
            const char *st = p;
            while (isalnum((unsigned char)*p) || *p == '_') p++;
            size_t n = (size_t)(p - st);
            char name[128];
            if (n >= sizeof name) {
                m4_put(m, st, n);
                continue;
            }
            memcpy(name, st, n);
            name[n] = '\0';
            if (strcmp(name, "dnl") == 0) {
                const char *e = strchr(p, '\n');
                p = e ? e + 1 : p + strlen(p);
                continue;
            }
            if ((strcmp(name, "define") == 0 || strcmp(name, "include") == 0) && *p == '(') {
                static char args[9][4096];
                int na;
                const char *after = m4_args(p, args, &na);
                if (!after) {
                    snprintf(m->err, sizeof m->err, "m4: unterminated %s(", name);
                    return -1;
                }
                p = after;
                if (name[0] == 'd') {
                    if (m->ndefs >= 256) {
                        snprintf(m->err, sizeof m->err, "m4: too many definitions");
                        return -1;
                    }
                    int slot = m->ndefs;
                    for (int k = 0; k < m->ndefs; k++)
                        if (strcmp(m->defs[k].name, args[0]) == 0) slot = k;
                    if (slot == m->ndefs) m->ndefs++;
                    else free(m->defs[slot].text);
                    snprintf(m->defs[slot].name, sizeof m->defs[slot].name, "%s", args[0]);
                    m->defs[slot].text = strdup(na > 1 ? args[1] : "");
                } else {
                    char *inc = slurp(args[0]);
                    if (!inc && m->dir[0]) {
                        char path[2200];
                        snprintf(path, sizeof path, "%s/%s", m->dir, args[0]);
                        inc = slurp(path);
                    }
                    if (!inc) {
                        snprintf(m->err, sizeof m->err, "m4: cannot open include file '%s'", args[0]);
                        return -1;
                    }
                    int r = m4_run(m, inc, depth + 1);
                    free(inc);
                    if (r) return r;
                }
                continue;
            }
            int hit = -1;
            for (int k = 0; k < m->ndefs; k++)
                if (strcmp(m->defs[k].name, name) == 0) hit = k;
            if (hit >= 0) {
                int r = m4_run(m, m->defs[hit].text, depth + 1);
                if (r) return r;
            } else {
                m4_put(m, st, n);
            }
            continue;
        }
        m4_put(m, p, 1); /* digits and punctuation are single characters to m4 */
        p++;
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * line reader and tags (sppde_parser.c:63-317)
 * ---------------------------------------------------------------------- */
typedef struct {
    char *buf;     /* pre-processed text */
    char *cur;     /* next unread position */
    char line[4096];
    char err[512];
} lreader;

static int cleanline(char *q) { /* cleanstr: cut at '\n' or '#', report non-blank */
    int text = 0;
    while (*q) {
        if (*q == '\n' || *q == '#') {
            *q = '\0';
            break;
        }
        if (*q != ' ' && *q != '\t') text = 1;
        q++;
    }
    return text;
}

typedef struct {
    char *buf;
    size_t n, cap;
} sbuf;

static void sb_add(sbuf *b, const char *s, size_t n) {
    if (b->n + n >= b->cap) n = b->cap - 1 - b->n; /* lines are bounded: truncate as SW2's fixed buffers would */
    memcpy(b->buf + b->n, s, n);
    b->n += n;
    b->buf[b->n] = '\0';
}

/* The value of every <expr> on a line, written back in its place with
 * %.15e (what SW2's reader does before it scans a line,
 * sppde_parser.c:84-143).  The line is taken as a run of segments:
 *   text        copied as it stands (a backslash before any other character
 *               stays; a backslash that ends the line is dropped);
 *   <expr>      evaluated and replaced by its value; no backslash inside;
 *   \< or \>   the bracket itself, and from there on the line is text in
 *               which only these two escapes are still read (the bracket
 *               kept, the backslash dropped): SW2's way of writing a
 *               literal '<' or '>'.
 * A '<' without its '>', or a '>' outside any <expr>, is an error. */
static int subst_exprs(lreader *r, char *line) {
    char outb[4096];
    sbuf out = {outb, 0, sizeof outb};
    outb[0] = '\0';
    int literal = 0; /* after the first escaped bracket */
    const char *q = line;
    while (*q) {
        const int esc = q[0] == '\\' && (q[1] == '<' || q[1] == '>');
        if (esc) {
            sb_add(&out, q + 1, 1);
            q += 2;
            literal = 1;
            continue;
        }
        if (q[0] == '\\' && q[1] == '\0') break;
        if (!literal && *q == '<') {
            const char *close = strchr(q + 1, '>');
            if (!close) {
                snprintf(r->err, sizeof r->err, "missing '>' while evaluating '%s'", line);
                return -1;
            }
            char ex[4096];
            size_t len = (size_t)(close - (q + 1));
            if (len >= sizeof ex) len = sizeof ex - 1;
            memcpy(ex, q + 1, len);
            ex[len] = '\0';
            int bad = strchr(ex, '\\') != NULL;
            double v = bad ? NAN : sw2_eval_expr(ex, &bad);
            if (bad) {
                snprintf(r->err, sizeof r->err, "Expression='%s' can not be evaluated", ex);
                return -1;
            }
            char num[64];
            int k = snprintf(num, sizeof num, "%.15e", v);
            sb_add(&out, num, (size_t)k);
            q = close + 1;
            continue;
        }
        if (!literal && *q == '>') {
            snprintf(r->err, sizeof r->err, "'>' without its '<' in '%s'", line);
            return -1;
        }
        sb_add(&out, q, 1);
        q++;
    }
    memcpy(line, outb, out.n + 1);
    return 0;
}

static char *nextline(lreader *r) {
    while (r->cur && *r->cur) {
        char *e = strchr(r->cur, '\n');
        size_t n = e ? (size_t)(e - r->cur) : strlen(r->cur);
        if (n > sizeof r->line - 1) n = sizeof r->line - 1;
        memcpy(r->line, r->cur, n);
        r->line[n] = '\0';
        r->cur = e ? e + 1 : r->cur + strlen(r->cur);
        if (cleanline(r->line)) {
            if (subst_exprs(r, r->line)) return NULL;
            return r->line;
        }
    }
    snprintf(r->err, sizeof r->err, "Unexpected end of file");
    return NULL;
}

/* A line as SW2's tag readers take it (sppde_parser.c:161-317): the next
 * line with text once comments are cut and <expr> values written in, split
 * into words at blanks and tabs.  With a tag, the first word must be that
 * tag; the words after it are the values. */
typedef struct {
    char *w[128];
    int n;
} words;

static int next_words(lreader *r, const char *tag, words *W) {
    char *l = nextline(r);
    if (!l) return -1;
    char *save = NULL;
    W->n = 0;
    for (char *t = strtok_r(l, " \t", &save); t && W->n < 128; t = strtok_r(NULL, " \t", &save)) W->w[W->n++] = t;
    if (!tag[0]) return 0;
    if (W->n == 0 || strcmp(W->w[0], tag) != 0) {
        snprintf(r->err, sizeof r->err, "expecting '%s' got '%s'", tag, W->n ? W->w[0] : "");
        return -1;
    }
    memmove(W->w, W->w + 1, sizeof(char *) * (size_t)(W->n - 1));
    W->n--;
    return 0;
}

/* the number a word starts with (what %le reads from it), or -1 */
static int word_number(const char *w, double *v) {
    char *end = NULL;
    *v = strtod(w, &end);
    return end == w ? -1 : 0;
}

static int getdouble(lreader *r, const char *tag, double *out) {
    words W;
    if (next_words(r, tag, &W)) return -1;
    if (W.n == 0 || word_number(W.w[0], out)) {
        snprintf(r->err, sizeof r->err, "reading tag '%s' expected valid double and found '%s'", tag, W.n ? W.w[0] : "");
        return -1;
    }
    return 0;
}

/* an integer tag: the value must have no fractional part (SW2 reads it as
 * a double first, so "1e3" is 1000) */
static int getint(lreader *r, const char *tag, int *out) {
    words W;
    double v;
    if (next_words(r, tag, &W)) return -1;
    if (W.n == 0 || word_number(W.w[0], &v)) {
        snprintf(r->err, sizeof r->err, "expecting valid integer and found '%s'", W.n ? W.w[0] : "");
        return -1;
    }
    if (v != trunc(v)) {
        snprintf(r->err, sizeof r->err, "expecting an integer and found '%s'", W.w[0]);
        return -1;
    }
    *out = (int)v;
    return 0;
}

static int getndoubles(lreader *r, const char *tag, int nd, double *d) {
    words W;
    if (next_words(r, tag, &W)) return -1;
    for (int k = 0; k < nd; k++) {
        if (k >= W.n) {
            snprintf(r->err, sizeof r->err, "expecting more doubles nd=%d i=%d", nd, k + 1);
            return -1;
        }
        if (word_number(W.w[k], d + k)) {
            snprintf(r->err, sizeof r->err, "expecting one double and found=<%s>", W.w[k]);
            return -1;
        }
    }
    return 0;
}

/* a table: its row count, its column count, then one line per row */
static int gettable(lreader *r, int *nr, int *nc, double **out) {
    if (getint(r, "", nr) || getint(r, "", nc)) return -1;
    if (*nc > 100 || *nr < 0 || *nc < 0) {
        snprintf(r->err, sizeof r->err, "bad table shape %d x %d", *nr, *nc);
        return -1;
    }
    double *t = (double *)calloc((size_t)(*nr) * (*nc) + 1, sizeof(double));
    for (int row = 0; row < *nr; row++)
        if (getndoubles(r, "", *nc, t + (size_t)row * (*nc))) {
            free(t);
            return -1;
        }
    *out = t;
    return 0;
}

int sw2_preprocess(const char *fname, char **out, char *err, int errlen) {
    m4state *m = (m4state *)calloc(1, sizeof(m4state));
    char *txt = slurp(fname);
    if (!txt) {
        snprintf(err, errlen, "could not open '%s'", fname);
        free(m);
        return -1;
    }
    const char *slash = strrchr(fname, '/');
    if (slash) {
        size_t n = (size_t)(slash - fname);
        if (n >= sizeof m->dir) n = sizeof m->dir - 1;
        memcpy(m->dir, fname, n);
        m->dir[n] = '\0';
    }
    int rc = m4_run(m, txt, 0);
    free(txt);
    for (int k = 0; k < m->ndefs; k++) free(m->defs[k].text);
    if (rc) {
        snprintf(err, errlen, "%s", m->err);
        free(m->out);
        free(m);
        return -1;
    }
    *out = m->out ? m->out : strdup("");
    free(m);
    return 0;
}

#define TRY(x) do { if ((x)) goto fail; } while (0)

int sw2_read_case(const char *fname, sw2_params *P) {
    memset(P, 0, sizeof *P);
    lreader r;
    memset(&r, 0, sizeof r);
    char perr[512];
    if (sw2_preprocess(fname, &r.buf, perr, sizeof perr)) {
        sw2_set_error("%s", perr);
        return -1;
    }
    r.cur = r.buf;
    double *vort = NULL, *gauss = NULL, *wind = NULL, *reft = NULL;
    int nr, nc;
    TRY(getint(&r, "sch", &P->sch));
    TRY(getint(&r, "polar", &P->polar));
    if (P->polar != 0) {
        snprintf(r.err, sizeof r.err, "polar=%d: only the channel geometry (polar 0) runs on the GPU path", P->polar);
        goto fail;
    }
    TRY(getdouble(&r, "lon0", &P->lon0_deg));
    TRY(getdouble(&r, "lon1", &P->lon1_deg));
    TRY(getdouble(&r, "lat0", &P->lat0_deg));
    TRY(getdouble(&r, "lat1", &P->lat1_deg));
    TRY(getint(&r, "nx", &P->nx));
    TRY(getint(&r, "ny", &P->ny));
    TRY(getdouble(&r, "t0", &P->t0));
    TRY(getdouble(&r, "t1", &P->t1));
    TRY(getdouble(&r, "Dt", &P->Dt));
    TRY(getdouble(&r, "CFL", &P->CFL));
    TRY(getdouble(&r, "Tau", &P->Tau));
    TRY(getndoubles(&r, "Hypernu", 3, P->hypernu));
    TRY(getint(&r, "IteInfo", &P->IteInfo));
    TRY(getint(&r, "SaveEvery", &P->SaveEvery));
    TRY(getint(&r, "LoadFrom", &P->LoadFrom));
    TRY(getint(&r, "reftype", &P->reftype));
    if (P->reftype == 1) {
        words W;
        TRY(next_words(&r, "refname", &W));
        TRY(gettable(&r, &nr, &nc, &reft));
        free(reft);
        reft = NULL;
    } else if (P->reftype != 0) {
        snprintf(r.err, sizeof r.err, "wrong reftype=%d", P->reftype);
        goto fail;
    }
    TRY(getdouble(&r, "Omega", &P->Omega));
    TRY(getdouble(&r, "rE", &P->rE));
    TRY(getdouble(&r, "rP", &P->rP));
    TRY(getdouble(&r, "gE", &P->gE));
    TRY(getdouble(&r, "depth", &P->depth));
    TRY(getint(&r, "geoeq", &P->geoeq));
    TRY(gettable(&r, &nr, &nc, &vort));
    if (nc != N_COLS_VORTICES) {
        snprintf(r.err, sizeof r.err, "Each SW vortex should be defined with %d columns", N_COLS_VORTICES);
        goto fail;
    }
    P->num_vortices = nr;
    P->vortices = vort;
    TRY(gettable(&r, &nr, &nc, &gauss));
    if (nc != N_COLS_GAUSS) {
        snprintf(r.err, sizeof r.err, "Each SW perturbation should be defined with %d columns", N_COLS_GAUSS);
        goto fail;
    }
    P->num_Gaussians = nr;
    P->Gaussians = gauss;
    TRY(getdouble(&r, "TracerDis", &P->TracerDis));
    TRY(gettable(&r, &nr, &nc, &wind));
    if (nc != 2) {
        snprintf(r.err, sizeof r.err, "wind table must have 2 columns (lat_deg u), has %d", nc);
        goto fail;
    }
    P->wind_rows = nr;
    P->wind = wind;
    free(r.buf);
    return 0;
fail:
    sw2_set_error("case file %s: %s", fname, r.err);
    free(r.buf);
    free(vort);
    free(gauss);
    free(wind);
    P->vortices = P->Gaussians = P->wind = NULL;
    return -1;
}

void sw2_free_params(sw2_params *P) {
    free((void *)P->vortices);
    free((void *)P->Gaussians);
    free((void *)P->wind);
    P->vortices = P->Gaussians = P->wind = NULL;
}
