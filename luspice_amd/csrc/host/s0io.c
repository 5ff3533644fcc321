/* s0io.c -- s0 text files: matrix, right-hand side and .flt pin list.
 *
 * Format, as the reference's readers accept it (src/spmv.cu:42-145,
 * src/mp/mp_lu_bench.cu:46-111):
 *   matrix  first line "n <n>"; entries "A <i> <j> <v>", 1-based; every other
 *           line is ignored
 *   rhs     entries "RHS <i> <v>", 1-based
 *   flt     pins "F <row>"
 * What callers observe, kept as the reference has it:
 *   read_s0_to_csr  rows in file order inside each row; an entry with
 *                   i outside [1, n] or j < 1 is dropped (j > n is kept);
 *                   a file with no "A" line at all is an error
 *   read_rhs        the first line is skipped; x[i-1] = v for 1 <= i <= n,
 *                   other x entries untouched; 0 when at least one was read
 *   lsp_s0_read     the COO in file order (1-based, unchecked), b[1..n] from
 *                   the RHS file (any line, 1 <= i <= n); an entry that is
 *                   exactly zero or missing becomes 1.0 (mp_lu_bench.cu:83)
 *   lsp_flt_pin     each pin adds `pin` to the first diagonal entry of its
 *                   row in COO order, or appends (row, row, pin) when the row
 *                   has none (an appended entry takes later pins of its row);
 *                   nothing happens when pin <= 0 or the file does not open
 *
 * Design: a file is read whole into one buffer and cut into lines in place
 * (each newline becomes a terminator); one record parser classifies a line
 * by its tag and reads its numbers with strtol / strtod.  Every reader is a
 * loop over records.  Arrays are sized from the line count, an upper bound on
 * the records, so a file is scanned once and nothing is regrown.
 */
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spmv.h"
#include "s0io.h"

/* ---- whole-file text and its lines ------------------------------------- */
typedef struct {
    char *text;     /* NUL-terminated copy of the file */
    char *pos;      /* start of the next line */
    char *end;
} TextFile;

static int text_load(const char *path, TextFile *t)
{
    memset(t, 0, sizeof(*t));
    if (!path) return 1;
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    size_t cap = (size_t)1 << 16, len = 0;
    char *buf = (char *)malloc(cap + 1);
    while (buf) {
        len += fread(buf + len, 1, cap - len, f);
        if (len < cap) break;
        char *grown = (char *)realloc(buf, 2 * cap + 1);
        if (!grown) { free(buf); buf = NULL; break; }
        buf = grown;
        cap *= 2;
    }
    const int bad = ferror(f);
    fclose(f);
    if (!buf || bad) { free(buf); return 1; }
    buf[len] = '\0';
    t->text = t->pos = buf;
    t->end = buf + len;
    return 0;
}

static void text_free(TextFile *t)
{
    free(t->text);
    memset(t, 0, sizeof(*t));
}

/* Lines left from the cursor on (an upper bound on the records in them). */
static size_t lines_left(const TextFile *t)
{
    size_t n = 1;
    for (const char *p = t->pos; (p = (const char *)memchr(p, '\n', (size_t)(t->end - p))) != NULL; p++) n++;
    return n;
}

/* The next line, its newline replaced by a terminator; NULL at the end. */
static char *next_line(TextFile *t)
{
    if (t->pos >= t->end) return NULL;
    char *ln = t->pos;
    char *nl = (char *)memchr(ln, '\n', (size_t)(t->end - ln));
    if (nl) {
        *nl = '\0';
        t->pos = nl + 1;
    } else {
        t->pos = t->end;
    }
    return ln;
}

/* ---- records ------------------------------------------------------------- */
enum { REC_NONE = 0, REC_HEADER, REC_ENTRY, REC_RHS, REC_PIN };

typedef struct {
    int i, j;       /* header: i = n; entry: (i, j); rhs: i; pin: i = row */
    double v;
} Record;

/* An int field (leading blanks, sign); false when absent or out of range. */
static int take_int(const char **s, int *out)
{
    char *e;
    errno = 0;
    const long v = strtol(*s, &e, 10);
    if (e == *s || errno == ERANGE || v < INT_MIN || v > INT_MAX) return 0;
    *s = e;
    *out = (int)v;
    return 1;
}

static int take_real(const char **s, double *out)
{
    char *e;
    const double v = strtod(*s, &e);
    if (e == *s) return 0;
    *s = e;
    *out = v;
    return 1;
}

/* Kind of the line and its fields; the tag must open the line. */
static int parse_record(const char *ln, Record *r)
{
    const char *s = ln;
    switch (ln[0]) {
    case 'n':
        s++;
        return take_int(&s, &r->i) ? REC_HEADER : REC_NONE;
    case 'A':
        s++;
        return take_int(&s, &r->i) && take_int(&s, &r->j) && take_real(&s, &r->v) ? REC_ENTRY : REC_NONE;
    case 'R':
        if (strncmp(ln, "RHS", 3) != 0) return REC_NONE;
        s += 3;
        return take_int(&s, &r->i) && take_real(&s, &r->v) ? REC_RHS : REC_NONE;
    case 'F':
        s++;
        return take_int(&s, &r->i) ? REC_PIN : REC_NONE;
    default:
        return REC_NONE;
    }
}

/* The order on the first line, or 0 when the line is not a valid header. */
static int read_header(TextFile *t)
{
    const char *ln = next_line(t);
    Record r;
    if (!ln || parse_record(ln, &r) != REC_HEADER || r.i <= 0) return 0;
    return r.i;
}

/* ---- spmv.h readers ------------------------------------------------------ */
int read_s0_to_csr(const char *path, CSRMatrix *A)
{
    TextFile t;
    if (text_load(path, &t)) {
        fprintf(stderr, "s0: cannot read matrix file %s\n", path ? path : "(null)");
        return 1;
    }
    const int n = read_header(&t);
    if (n <= 0) { text_free(&t); return 1; }
    const size_t cap = lines_left(&t);
    int *er = (int *)malloc(cap * sizeof(int));          /* entry rows, 0-based */
    int *ec = (int *)malloc(cap * sizeof(int));
    double *ev = (double *)malloc(cap * sizeof(double));
    int *rowptr = (int *)calloc((size_t)n + 1, sizeof(int));
    if (!er || !ec || !ev || !rowptr) {
        free(er); free(ec); free(ev); free(rowptr); text_free(&t);
        return 1;
    }
    long long tagged = 0, k = 0;
    const char *ln;
    while ((ln = next_line(&t)) != NULL) {
        if (ln[0] != 'A') continue;
        tagged++;
        Record r;
        if (parse_record(ln, &r) != REC_ENTRY || r.i < 1 || r.i > n || r.j < 1) continue;
        er[k] = r.i - 1;
        ec[k] = r.j - 1;
        ev[k] = r.v;
        rowptr[r.i]++;                                   /* count of row i-1 */
        k++;
    }
    text_free(&t);
    if (tagged == 0) {
        free(er); free(ec); free(ev); free(rowptr);
        return 1;
    }
    for (int i = 0; i < n; i++) rowptr[i + 1] += rowptr[i];
    /* place entries row by row in file order: fill[i] is row i's next slot */
    int *fill = (int *)malloc((size_t)n * sizeof(int));
    int *colidx = (int *)malloc((size_t)(k ? k : 1) * sizeof(int));
    double *values = (double *)malloc((size_t)(k ? k : 1) * sizeof(double));
    if (!fill || !colidx || !values) {
        free(fill); free(colidx); free(values); free(er); free(ec); free(ev); free(rowptr);
        return 1;
    }
    memcpy(fill, rowptr, (size_t)n * sizeof(int));
    for (long long e = 0; e < k; e++) {
        const int slot = fill[er[e]]++;
        colidx[slot] = ec[e];
        values[slot] = ev[e];
    }
    free(fill); free(er); free(ec); free(ev);
    A->n = n;
    A->nnz = k;
    A->rowptr = rowptr;
    A->colidx = colidx;
    A->values = values;
    return 0;
}

void csr_free(CSRMatrix *A)
{
    if (!A) return;
    free(A->rowptr);
    free(A->colidx);
    free(A->values);
    memset(A, 0, sizeof(*A));
}

int read_rhs(const char *path, int n, double *x)
{
    TextFile t;
    if (text_load(path, &t)) {
        fprintf(stderr, "s0: cannot read right-hand-side file %s\n", path ? path : "(null)");
        return 1;
    }
    int got = 0;
    const char *ln = next_line(&t);                     /* the header line */
    while (ln && (ln = next_line(&t)) != NULL) {
        Record r;
        if (parse_record(ln, &r) == REC_RHS && r.i >= 1 && r.i <= n) {
            x[r.i - 1] = r.v;
            got++;
        }
    }
    text_free(&t);
    return got > 0 ? 0 : 1;
}

/* ---- the CLI reader ------------------------------------------------------ */
int lsp_s0_read(const char *mtx, const char *rhs, lsp_s0 *M)
{
    memset(M, 0, sizeof(*M));
    TextFile t;
    if (text_load(mtx, &t)) {
        fprintf(stderr, "s0: cannot read matrix file %s\n", mtx ? mtx : "(null)");
        return 1;
    }
    const int n = read_header(&t);
    if (n <= 0) { text_free(&t); return 1; }
    const size_t cap = lines_left(&t);
    M->n = n;
    M->row = (int *)malloc(cap * sizeof(int));
    M->col = (int *)malloc(cap * sizeof(int));
    M->val = (double *)malloc(cap * sizeof(double));
    M->b = (double *)calloc((size_t)n + 1, sizeof(double));
    if (!M->row || !M->col || !M->val || !M->b) {
        text_free(&t);
        lsp_s0_free(M);
        return 1;
    }
    const char *ln;
    while ((ln = next_line(&t)) != NULL) {
        Record r;
        if (ln[0] != 'A' || parse_record(ln, &r) != REC_ENTRY) continue;
        M->row[M->nnz] = r.i;
        M->col[M->nnz] = r.j;
        M->val[M->nnz] = r.v;
        M->nnz++;
    }
    text_free(&t);
    if (rhs && text_load(rhs, &t) == 0) {
        while ((ln = next_line(&t)) != NULL) {
            Record r;
            if (parse_record(ln, &r) == REC_RHS && r.i >= 1 && r.i <= n) M->b[r.i] = r.v;
        }
        text_free(&t);
    }
    for (int i = 1; i <= n; i++)
        if (M->b[i] == 0.0) M->b[i] = 1.0;
    return 0;
}

void lsp_s0_free(lsp_s0 *M)
{
    free(M->row);
    free(M->col);
    free(M->val);
    free(M->b);
    memset(M, 0, sizeof(*M));
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Index of row r in the sorted distinct rows u[0..nu), or -1. */
static long find_row(const int *u, long nu, int r)
{
    long lo = 0, hi = nu;
    while (lo < hi) {
        const long mid = lo + (hi - lo) / 2;
        if (u[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    return lo < nu && u[lo] == r ? lo : -1;
}

int lsp_flt_pin(lsp_s0 *M, const char *flt, double pin)
{
    if (!flt || !(pin > 0)) return 0;
    TextFile t;
    if (text_load(flt, &t)) return 0;
    /* the pinned rows, in file order */
    int *pins = (int *)malloc(lines_left(&t) * sizeof(int));
    long np = 0;
    const char *ln;
    if (pins) {
        while ((ln = next_line(&t)) != NULL) {
            Record r;
            if (parse_record(ln, &r) == REC_PIN) pins[np++] = r.i;
        }
    }
    text_free(&t);
    if (np == 0) { free(pins); return 0; }
    /* distinct pinned rows, and the first diagonal entry of each in COO order */
    int *rows = (int *)malloc((size_t)np * sizeof(int));
    long *at = (long *)malloc((size_t)np * sizeof(long));
    int *nr = (int *)realloc(M->row, (size_t)(M->nnz + np) * sizeof(int));
    if (nr) M->row = nr;
    int *nc = (int *)realloc(M->col, (size_t)(M->nnz + np) * sizeof(int));
    if (nc) M->col = nc;
    double *nv = (double *)realloc(M->val, (size_t)(M->nnz + np) * sizeof(double));
    if (nv) M->val = nv;
    if (!rows || !at || !nr || !nc || !nv) {
        free(pins); free(rows); free(at);
        return 0;
    }
    memcpy(rows, pins, (size_t)np * sizeof(int));
    qsort(rows, (size_t)np, sizeof(int), cmp_int);
    long nu = 0;
    for (long p = 0; p < np; p++)
        if (nu == 0 || rows[nu - 1] != rows[p]) rows[nu++] = rows[p];
    for (long u = 0; u < nu; u++) at[u] = -1;
    for (long e = 0; e < M->nnz; e++) {
        if (M->row[e] != M->col[e]) continue;
        const long u = find_row(rows, nu, M->row[e]);
        if (u >= 0 && at[u] < 0) at[u] = e;
    }
    int appended = 0;
    for (long p = 0; p < np; p++) {
        const long u = find_row(rows, nu, pins[p]);
        if (at[u] >= 0) {
            M->val[at[u]] += pin;
        } else {
            at[u] = M->nnz;
            M->row[M->nnz] = pins[p];
            M->col[M->nnz] = pins[p];
            M->val[M->nnz] = pin;
            M->nnz++;
            appended++;
        }
    }
    free(pins); free(rows); free(at);
    return appended;
}

/* ---- writers ------------------------------------------------------------- */
int lsp_s0_write(const char *path, int n, long nnz, const int *row0, const int *col0,
                 const double *val)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, "n %d\n", n);
    for (long e = 0; e < nnz; e++) fprintf(f, "A %d %d %.17g\n", row0[e] + 1, col0[e] + 1, val[e]);
    return fclose(f) ? 1 : 0;
}

int lsp_rhs_write(const char *path, int n, const double *b)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, "n %d\n", n);
    for (int i = 0; i < n; i++) fprintf(f, "RHS %d %.17g\n", i + 1, b[i]);
    return fclose(f) ? 1 : 0;
}
