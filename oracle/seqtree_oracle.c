/* seqtree_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithms for the all-vs-all scoring path
 * (seqtree 0.6.0). It is the checker: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and nothing in the shipped library calls it.
 *
 * Each function cites the reference lines it restates. The matrix drivers keep the
 * reference's threading (a std::thread-style worker pool taking query rows from an atomic
 * counter, reference src/gapblock.cpp:105-131, src/pairwise.cpp:298-331,
 * src/distance.cpp:36-62) so the file doubles as a "port" CPU baseline.
 *
 * Pinning (tests/test_oracle.py): every function is checked against the compiled reference
 * (oracle/_ref) where it builds, against the reference's own known answers, and against
 * independent brute-force formulations (enumeration of all gap-block layouts, a full
 * three-matrix Gotoh, the textbook Levenshtein DP).
 *
 * Return codes: 0 ok, 1 invalid argument (Hamming length mismatch), 2 overflow.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------------------
 * Substitution matrices (reference src/substitution_matrix.cpp:29-50, 76-86)
 * ------------------------------------------------------------------------------------- */

/* pen[a][b] = max(0, s[a][a] + s[b][b] - 2 s[a][b]) in wrapping int32 arithmetic. */
void or_pen_from_sim(int32_t size, const int32_t* sim, int32_t* pen) {
    for (int32_t a = 0; a < size; ++a)
        for (int32_t b = 0; b < size; ++b) {
            uint32_t v = (uint32_t)sim[a * size + a] + (uint32_t)sim[b * size + b] -
                         2u * (uint32_t)sim[a * size + b];
            int32_t p = (int32_t)v;
            pen[a * size + b] = p < 0 ? 0 : p;
        }
}

static int cmp_i32(const void* x, const void* y) {
    int32_t a = *(const int32_t*)x, b = *(const int32_t*)y;
    return (a > b) - (a < b);
}

/* Median of the off-diagonal penalties: the element of rank count/2 (nth_element). */
int32_t or_scale(int32_t size, const int32_t* pen) {
    if (size < 2) return 0;
    int32_t n = size * (size - 1), k = 0;
    int32_t* off = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t a = 0; a < size; ++a)
        for (int32_t b = 0; b < size; ++b)
            if (a != b) off[k++] = pen[a * size + b];
    qsort(off, (size_t)n, sizeof(int32_t), cmp_i32);
    int32_t v = off[n / 2];
    free(off);
    return v;
}

/* ---------------------------------------------------------------------------------------
 * Codec (reference src/codec.cpp:9-35): returns the first invalid position or -1.
 * alphabet: 0 amino acid, 1 nucleotide, 2 IUPAC nucleotide.
 * ------------------------------------------------------------------------------------- */
int64_t or_encode(int32_t alphabet, const uint8_t* in, uint64_t n, uint8_t* out) {
    static const char* syms[3] = {"ARNDCQEGHILKMFPSTWYVBZX*", "ACGT", "ACGTRYSWKMBDHVN"};
    uint8_t table[256];
    memset(table, 0xFF, sizeof(table));
    const char* s = syms[alphabet < 0 || alphabet > 2 ? 0 : alphabet];
    for (uint8_t c = 0; s[c]; ++c) {
        uint8_t up = (uint8_t)s[c];
        table[up] = c;
        if (up >= 'A' && up <= 'Z') table[up - 'A' + 'a'] = c;
    }
    for (uint64_t i = 0; i < n; ++i) {
        out[i] = table[in[i]];
        if (out[i] == 0xFF) return (int64_t)i;
    }
    return -1;
}

/* ---------------------------------------------------------------------------------------
 * Gap-block cell (reference src/gapblock.cpp:39-62): suffix sum on diagonal d, then a
 * forward prefix sweep over the L+1 block positions with the prior, int32 wrapping as the
 * reference's int32 arithmetic does in practice.
 * ------------------------------------------------------------------------------------- */
int32_t or_gapblock_cell(const uint8_t* q, uint32_t m, const uint8_t* r, uint32_t n,
                         const int32_t* pen, int32_t A, int32_t gap_open, int32_t gap_extend,
                         const int32_t* prior_row, uint32_t* suf /* >= min(m,n)+1 */) {
    uint32_t L = m < n ? m : n;
    uint32_t d = m > n ? m - n : n - m;
    int q_longer = m >= n;
    suf[L] = 0;
    for (uint32_t j = L; j-- > 0;) {
        uint8_t a = q_longer ? q[j + d] : q[j];
        uint8_t b = q_longer ? r[j] : r[j + d];
        suf[j] = suf[j + 1] + (uint32_t)pen[(size_t)a * (size_t)A + b];
    }
    if (d == 0) return (int32_t)suf[0];
    uint32_t pre = 0;
    int32_t best = INT32_MAX;
    for (uint32_t i = 0; i <= L; ++i) {
        int32_t cand = (int32_t)(pre + suf[i] + (prior_row ? (uint32_t)prior_row[i] : 0u));
        if (cand < best) best = cand;
        if (i < L) pre += (uint32_t)pen[(size_t)q[i] * (size_t)A + r[i]];
    }
    return (int32_t)((uint32_t)best + (uint32_t)gap_open + (d - 1) * (uint32_t)gap_extend);
}

/* ---------------------------------------------------------------------------------------
 * Gotoh score (reference src/pairwise.cpp:71-125): int64 two-row DP.
 * ------------------------------------------------------------------------------------- */
static int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

static int64_t gap_run(int64_t L, int32_t open, int32_t extend) {
    return L <= 0 ? 0 : (int64_t)open + (L - 1) * (int64_t)extend;
}

int64_t or_gotoh(const uint8_t* q, uint64_t m, const uint8_t* r, uint64_t n, const int32_t* sim,
                 int32_t A, int32_t local, int32_t go, int32_t ge) {
    const int64_t NEG = INT64_MIN / 4;
    if (m == 0 || n == 0) {
        if (local) return 0;
        return -gap_run((int64_t)(m ? m : n), go, ge);
    }
    int64_t* buf = (int64_t*)malloc(sizeof(int64_t) * 6 * (size_t)(n + 1));
    int64_t *Mp = buf, *Xp = Mp + n + 1, *Yp = Xp + n + 1;
    int64_t *Mc = Yp + n + 1, *Xc = Mc + n + 1, *Yc = Xc + n + 1;
    Mp[0] = 0;
    Xp[0] = NEG;
    Yp[0] = NEG;
    for (uint64_t j = 1; j <= n; ++j) {
        Mp[j] = local ? 0 : NEG;
        Xp[j] = NEG;
        Yp[j] = local ? NEG : -gap_run((int64_t)j, go, ge);
    }
    int64_t best = local ? 0 : NEG;
    for (uint64_t i = 1; i <= m; ++i) {
        Mc[0] = local ? 0 : NEG;
        Xc[0] = local ? NEG : -gap_run((int64_t)i, go, ge);
        Yc[0] = NEG;
        for (uint64_t j = 1; j <= n; ++j) {
            int64_t s = sim[(size_t)q[i - 1] * (size_t)A + r[j - 1]];
            int64_t diag = max64(Mp[j - 1], max64(Xp[j - 1], Yp[j - 1]));
            int64_t mv = diag + s;
            if (local && mv < 0) mv = 0;
            Mc[j] = mv;
            Xc[j] = max64(Mp[j] - go, max64(Xp[j] - ge, Yp[j] - go));
            Yc[j] = max64(Mc[j - 1] - go, max64(Yc[j - 1] - ge, Xc[j - 1] - go));
            if (local && mv > best) best = mv;
        }
        int64_t* t;
        t = Mp; Mp = Mc; Mc = t;
        t = Xp; Xp = Xc; Xc = t;
        t = Yp; Yp = Yc; Yc = t;
    }
    int64_t res = local ? best : max64(Mp[n], max64(Xp[n], Yp[n]));
    free(buf);
    return res;
}

/* ---------------------------------------------------------------------------------------
 * Gotoh alignment with traceback (reference src/pairwise.cpp:141-267): full int64 M/X/Y
 * matrices and their back-pointers, the reference's argmax order (M, then X, then Y, each
 * replacing only on a strictly larger value), the first strict maximum in row-major order as
 * the local end, and its traceback rules. Writes the score, the ops in forward order ('M'
 * equal codes, 'S' substitution, 'D' query residue against a gap, 'I' reference residue
 * against a gap), their count and the first aligned residue of each side (the aligned
 * strings are those residues, decoded, in op order).
 * ------------------------------------------------------------------------------------- */
enum { OR_M = 0, OR_X = 1, OR_Y = 2, OR_STOP = 3 };

static int64_t or_pick(int64_t vm, int64_t vx, int64_t vy, uint8_t* who) {
    int64_t v = vm;
    *who = OR_M;
    if (vx > v) { v = vx; *who = OR_X; }
    if (vy > v) { v = vy; *who = OR_Y; }
    return v;
}

void or_align_pair(const uint8_t* q, uint64_t m, const uint8_t* r, uint64_t n,
                   const int32_t* sim, int32_t A, int32_t local, int32_t go, int32_t ge,
                   int64_t* score, uint8_t* ops, uint64_t* n_ops, uint64_t* q_begin,
                   uint64_t* r_begin) {
    const int64_t NEG = INT64_MIN / 4;
    *n_ops = 0;
    *q_begin = 0;
    *r_begin = 0;
    if (m == 0 || n == 0) {
        *score = local ? 0 : -gap_run((int64_t)(m ? m : n), go, ge);
        if (!local) {
            for (uint64_t i = 0; i < m; ++i) ops[(*n_ops)++] = 'D';
            for (uint64_t j = 0; j < n; ++j) ops[(*n_ops)++] = 'I';
        }
        return;
    }
    const uint64_t W = n + 1, sz = (m + 1) * W;
    int64_t* M = (int64_t*)malloc(sizeof(int64_t) * 3 * sz);
    int64_t *X = M + sz, *Y = X + sz;
    uint8_t* bM = (uint8_t*)malloc(3 * sz);
    uint8_t *bX = bM + sz, *bY = bX + sz;
    for (uint64_t k = 0; k < sz; ++k) {
        M[k] = X[k] = Y[k] = NEG;
        bM[k] = OR_STOP;
        bX[k] = bY[k] = OR_M;
    }
    M[0] = 0;
    for (uint64_t j = 1; j <= n; ++j) {
        M[j] = local ? 0 : NEG;
        Y[j] = local ? NEG : -gap_run((int64_t)j, go, ge);
        bY[j] = j > 1 ? OR_Y : OR_M;
    }
    for (uint64_t i = 1; i <= m; ++i) {
        M[i * W] = local ? 0 : NEG;
        X[i * W] = local ? NEG : -gap_run((int64_t)i, go, ge);
        bX[i * W] = i > 1 ? OR_X : OR_M;
    }
    int64_t best = local ? 0 : NEG;
    uint64_t bi = 0, bj = 0;
    for (uint64_t i = 1; i <= m; ++i)
        for (uint64_t j = 1; j <= n; ++j) {
            const uint64_t k = i * W + j, kd = (i - 1) * W + (j - 1), ku = (i - 1) * W + j,
                           kl = i * W + (j - 1);
            uint8_t who;
            int64_t mv = or_pick(M[kd], X[kd], Y[kd], &who) +
                         (int64_t)sim[(size_t)q[i - 1] * (size_t)A + r[j - 1]];
            if (local && mv < 0) { mv = 0; who = OR_STOP; }
            M[k] = mv;
            bM[k] = who;
            X[k] = or_pick(M[ku] - go, X[ku] - ge, Y[ku] - go, &bX[k]);
            Y[k] = or_pick(M[kl] - go, X[kl] - go, Y[kl] - ge, &bY[k]);
            if (local && mv > best) { best = mv; bi = i; bj = j; }
        }
    uint64_t i, j;
    uint8_t st;
    if (local) {
        *score = best;
        i = bi;
        j = bj;
        st = OR_M;
        if (best == 0) { free(M); free(bM); return; }
    } else {
        const uint64_t end = m * W + n;
        *score = or_pick(M[end], X[end], Y[end], &st);
        i = m;
        j = n;
    }
    uint64_t cnt = 0;  /* ops written backwards from the end of a buffer of m + n */
    uint8_t* rev = (uint8_t*)malloc(m + n);
    while (i > 0 || j > 0) {
        const uint64_t k = i * W + j;
        if (st == OR_M) {
            if (i == 0 || j == 0) break;
            if (local && M[k] == 0) break;
            rev[cnt++] = q[i - 1] == r[j - 1] ? 'M' : 'S';
            st = bM[k];
            --i;
            --j;
            if (st == OR_STOP) break;
        } else if (st == OR_X) {
            if (i == 0) break;
            rev[cnt++] = 'D';
            st = bX[k];
            --i;
        } else {
            if (j == 0) break;
            rev[cnt++] = 'I';
            st = bY[k];
            --j;
        }
    }
    for (uint64_t t = 0; t < cnt; ++t) ops[t] = rev[cnt - 1 - t];
    *n_ops = cnt;
    *q_begin = i;
    *r_begin = j;
    free(rev);
    free(M);
    free(bM);
}

/* ---------------------------------------------------------------------------------------
 * Penalty-space global alignment with traceback (reference src/index.cpp:105-179,
 * Index::align): minimises substitution penalty plus affine gap cost, a gap run of length L
 * costing go + (L - 1) ge and an X<->Y switch opening a new run. Three full int32 matrices
 * (M both residues, X query residue against a gap, Y reference residue against a gap),
 * every cell not reachable from (0, 0) held at kInf = INT32_MAX / 4, and per state the
 * predecessor state; the minimum over (M, X, Y) keeps the first of equal values (M, then X,
 * then Y, replacing only on a strictly smaller one). pen is the A x A penalty table indexed
 * [query code][reference code], or NULL for unit cost (0 equal codes, 1 otherwise). The
 * arithmetic wraps like the reference's int32 code; the walk never leaves the matrix (the
 * reference's walk does only for gap costs near kInf, where it reads outside its strings).
 * Writes the penalty score, the ops in forward order ('M' equal codes, 'S', 'D' query
 * residue against a gap, 'I' reference residue against a gap) and their count.
 * ------------------------------------------------------------------------------------- */
static int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

static int32_t or_least(int32_t vm, int32_t vx, int32_t vy, uint8_t* state) {
    int32_t v = vm;
    *state = OR_M;
    if (vx < v) { v = vx; *state = OR_X; }
    if (vy < v) { v = vy; *state = OR_Y; }
    return v;
}

void or_index_align(const uint8_t* q, uint64_t m, const uint8_t* r, uint64_t n,
                    const int32_t* pen, int32_t A, int32_t go, int32_t ge, int32_t* score,
                    uint8_t* ops, uint64_t* n_ops) {
    const int32_t INF = INT32_MAX / 4;
    const uint64_t W = n + 1, cells = (m + 1) * W;
    /* cell k of state s: val[3 k + s], from[3 k + s] */
    int32_t* val = (int32_t*)malloc(sizeof(int32_t) * 3 * cells);
    uint8_t* from = (uint8_t*)calloc(3 * cells, 1);
    for (uint64_t k = 0; k < 3 * cells; ++k) val[k] = INF;
    val[0] = 0;  /* M at (0, 0) */
    for (uint64_t i = 1; i <= m; ++i) {  /* column 0: a leading run of X */
        val[3 * (i * W) + OR_X] = wrap_add(go, (int32_t)((uint32_t)(i - 1) * (uint32_t)ge));
        from[3 * (i * W) + OR_X] = i > 1 ? OR_X : OR_M;
    }
    for (uint64_t j = 1; j <= n; ++j) {  /* row 0: a leading run of Y */
        val[3 * j + OR_Y] = wrap_add(go, (int32_t)((uint32_t)(j - 1) * (uint32_t)ge));
        from[3 * j + OR_Y] = j > 1 ? OR_Y : OR_M;
    }
    for (uint64_t i = 1; i <= m; ++i)
        for (uint64_t j = 1; j <= n; ++j) {
            const int32_t* d = val + 3 * ((i - 1) * W + (j - 1));
            const int32_t* u = val + 3 * ((i - 1) * W + j);
            const int32_t* l = val + 3 * (i * W + (j - 1));
            int32_t* c = val + 3 * (i * W + j);
            uint8_t* f = from + 3 * (i * W + j);
            const uint8_t a = q[i - 1], b = r[j - 1];
            const int32_t sub = pen ? pen[(size_t)a * (size_t)A + b] : (a != b);
            c[OR_M] = wrap_add(sub, or_least(d[OR_M], d[OR_X], d[OR_Y], &f[OR_M]));
            c[OR_X] = or_least(wrap_add(u[OR_M], go), wrap_add(u[OR_X], ge),
                               wrap_add(u[OR_Y], go), &f[OR_X]);
            c[OR_Y] = or_least(wrap_add(l[OR_M], go), wrap_add(l[OR_X], go),
                               wrap_add(l[OR_Y], ge), &f[OR_Y]);
        }
    uint8_t st;
    const int32_t* end = val + 3 * (m * W + n);
    *score = or_least(end[OR_M], end[OR_X], end[OR_Y], &st);
    uint64_t i = m, j = n, cnt = 0;
    uint8_t* rev = (uint8_t*)malloc(m + n + 1);
    while (i > 0 || j > 0) {
        const uint8_t* f = from + 3 * (i * W + j);
        if (st == OR_M) {
            if (i == 0 || j == 0) break;
            rev[cnt++] = q[i - 1] == r[j - 1] ? 'M' : 'S';
            st = f[OR_M];
            --i;
            --j;
        } else if (st == OR_X) {
            if (i == 0) break;
            rev[cnt++] = 'D';
            st = f[OR_X];
            --i;
        } else {
            if (j == 0) break;
            rev[cnt++] = 'I';
            st = f[OR_Y];
            --j;
        }
    }
    for (uint64_t t = 0; t < cnt; ++t) ops[t] = rev[cnt - 1 - t];
    *n_ops = cnt;
    free(rev);
    free(from);
    free(val);
}

/* ---------------------------------------------------------------------------------------
 * Unit-cost distances (reference src/distance.cpp:68-95).
 * ------------------------------------------------------------------------------------- */
int32_t or_hamming(const uint8_t* a, uint64_t m, const uint8_t* b, uint64_t n) {
    if (m != n) return -1;
    int32_t d = 0;
    for (uint64_t i = 0; i < m; ++i) d += a[i] != b[i];
    return d;
}

int32_t or_levenshtein(const uint8_t* a, uint64_t m, const uint8_t* b, uint64_t n) {
    if (n > m) {
        const uint8_t* t = a;
        a = b;
        b = t;
        uint64_t tm = m;
        m = n;
        n = tm;
    }
    if (n == 0) return (int32_t)m;
    int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(n + 1));
    int32_t* cur = prev + n + 1;
    for (uint64_t j = 0; j <= n; ++j) prev[j] = (int32_t)j;
    for (uint64_t i = 1; i <= m; ++i) {
        cur[0] = (int32_t)i;
        for (uint64_t j = 1; j <= n; ++j) {
            int32_t cost = a[i - 1] == b[j - 1] ? 0 : 1;
            int32_t v = prev[j] + 1;
            if (cur[j - 1] + 1 < v) v = cur[j - 1] + 1;
            if (prev[j - 1] + cost < v) v = prev[j - 1] + cost;
            cur[j] = v;
        }
        int32_t* t = prev;
        prev = cur;
        cur = t;
    }
    int32_t res = prev[n];
    free(prev < cur ? prev : cur);
    return res;
}

/* ---------------------------------------------------------------------------------------
 * Matrix drivers with the reference's row pool.
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int kind; /* 0 gapblock, 1 gotoh score, 2 gotoh dist, 3 hamming, 4 levenshtein */
    const uint8_t *q, *r;
    const uint64_t *qo, *ro;
    uint64_t N, K;
    const int32_t* table; /* pen (gapblock) or sim (gotoh) */
    int32_t A, local, go, ge;
    const int32_t* prior;
    uint64_t W1;
    uint32_t longest;
    const int64_t *qself, *rself;
    int32_t* out;
    atomic_ullong next;
    atomic_int err; /* 0, or the first error code seen */
} job_t;

static void* worker(void* arg) {
    job_t* j = (job_t*)arg;
    uint32_t* suf = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)j->longest + 1));
    for (;;) {
        uint64_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->N || atomic_load(&j->err)) break;
        const uint8_t* q = j->q + j->qo[i];
        uint64_t m = j->qo[i + 1] - j->qo[i];
        int32_t* row = j->out + i * j->K;
        for (uint64_t k = 0; k < j->K; ++k) {
            const uint8_t* r = j->r + j->ro[k];
            uint64_t n = j->ro[k + 1] - j->ro[k];
            if (j->kind == 0) {
                uint64_t M = m > n ? m : n, d = m > n ? m - n : n - m;
                const int32_t* prow = j->prior ? j->prior + (M * j->W1 + d) * j->W1 : NULL;
                row[k] = or_gapblock_cell(q, (uint32_t)m, r, (uint32_t)n, j->table, j->A, j->go,
                                          j->ge, prow, suf);
            } else if (j->kind == 1 || j->kind == 2) {
                int64_t s = or_gotoh(q, m, r, n, j->table, j->A, j->local, j->go, j->ge);
                int64_t v = j->kind == 2 ? j->qself[i] + j->rself[k] - 2 * s : s;
                if (v < INT32_MIN || v > INT32_MAX) {
                    int expected = 0;
                    atomic_compare_exchange_strong(&j->err, &expected, 2);
                    break;
                }
                row[k] = (int32_t)v;
            } else if (j->kind == 3) {
                int32_t h = or_hamming(q, m, r, n);
                if (h < 0) {
                    int expected = 0;
                    atomic_compare_exchange_strong(&j->err, &expected, 1);
                    break;
                }
                row[k] = h;
            } else {
                row[k] = or_levenshtein(q, m, r, n);
            }
        }
    }
    free(suf);
    return NULL;
}

static int run(job_t* j, int threads) {
    if (j->N == 0 || j->K == 0) return 0;
    if (threads <= 0) threads = 1;
    if ((uint64_t)threads > j->N) threads = (int)j->N;
    atomic_init(&j->next, 0);
    atomic_init(&j->err, 0);
    uint32_t longest = 0;
    for (uint64_t i = 0; i < j->N; ++i)
        if (j->qo[i + 1] - j->qo[i] > longest) longest = (uint32_t)(j->qo[i + 1] - j->qo[i]);
    for (uint64_t k = 0; k < j->K; ++k)
        if (j->ro[k + 1] - j->ro[k] > longest) longest = (uint32_t)(j->ro[k + 1] - j->ro[k]);
    j->longest = longest;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, j);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    return atomic_load(&j->err);
}

/* Reference src/gapblock.cpp:66-133 (after validation). prior: NULL or the (W+1)^3 cube. */
int or_gapblock_matrix(const uint8_t* q, const uint64_t* qo, uint64_t N, const uint8_t* r,
                       const uint64_t* ro, uint64_t K, const int32_t* pen, int32_t A, int32_t go,
                       int32_t ge, const int32_t* prior, uint32_t prior_width, int threads,
                       int32_t* out) {
    job_t j;
    memset(&j, 0, sizeof(j));
    j.kind = 0;
    j.q = q; j.qo = qo; j.N = N;
    j.r = r; j.ro = ro; j.K = K;
    j.table = pen; j.A = A; j.go = go; j.ge = ge;
    j.prior = prior;
    j.W1 = (uint64_t)prior_width + 1;
    j.out = out;
    return run(&j, threads);
}

/* Reference src/pairwise.cpp:273-334 (after validation): dist uses self scores computed
 * once per sequence, then d = s(a,a) + s(b,b) - 2 s(a,b), narrowed with overflow check. */
int or_align_matrix(const uint8_t* q, const uint64_t* qo, uint64_t N, const uint8_t* r,
                    const uint64_t* ro, uint64_t K, const int32_t* sim, int32_t A, int32_t local,
                    int32_t dist, int32_t go, int32_t ge, int threads, int32_t* out) {
    job_t j;
    memset(&j, 0, sizeof(j));
    j.kind = dist ? 2 : 1;
    j.q = q; j.qo = qo; j.N = N;
    j.r = r; j.ro = ro; j.K = K;
    j.table = sim; j.A = A; j.local = local; j.go = go; j.ge = ge;
    j.out = out;
    int64_t *qs = NULL, *rs = NULL;
    if (dist && N && K) {
        qs = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
        rs = (int64_t*)malloc(sizeof(int64_t) * (size_t)K);
        for (uint64_t i = 0; i < N; ++i)
            qs[i] = or_gotoh(q + qo[i], qo[i + 1] - qo[i], q + qo[i], qo[i + 1] - qo[i], sim, A,
                             local, go, ge);
        for (uint64_t k = 0; k < K; ++k)
            rs[k] = or_gotoh(r + ro[k], ro[k + 1] - ro[k], r + ro[k], ro[k + 1] - ro[k], sim, A,
                             local, go, ge);
    }
    j.qself = qs;
    j.rself = rs;
    int rc = run(&j, threads);
    free(qs);
    free(rs);
    return rc;
}

/* Reference src/distance.cpp:29-64, 97-109. which: 0 Hamming, 1 Levenshtein. */
int or_distance_matrix(int which, const uint8_t* a, const uint64_t* ao, uint64_t N,
                       const uint8_t* b, const uint64_t* bo, uint64_t K, int threads,
                       int32_t* out) {
    job_t j;
    memset(&j, 0, sizeof(j));
    j.kind = which == 0 ? 3 : 4;
    j.q = a; j.qo = ao; j.N = N;
    j.r = b; j.ro = bo; j.K = K;
    j.out = out;
    return run(&j, threads);
}
