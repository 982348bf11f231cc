/*
 * oracle.c -- CPU restatement of blastp two-hit word finding and ungapped
 * X-drop extension.  TEST INFRASTRUCTURE: linked only by the parity tests,
 * the smoke check and the benchmark's CPU-baseline leg, never by the product.
 * Function-by-function citations are in oracle.h.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))
#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* BLOSUM62 in ncbistdaa order.  The 20x20 core is the published matrix
 * (Henikoff & Henikoff 1992) in its usual ARNDCQEGHILKMFPSTWYV layout; the
 * ambiguity rows B, Z, J are the values the reference stages on its GPU
 * (gpu_blastp.c:46-75); X, U, O score -1 against everything but '*'; '*'
 * scores -4 against everything but itself (+1); the gap code scores -128. */
static const char kCore[] = "ARNDCQEGHILKMFPSTWYV";
static const signed char kCore20[20][20] = {
    /*       A   R   N   D   C   Q   E   G   H   I   L   K   M   F   P   S   T   W   Y   V */
    /*A*/ {  4, -1, -2, -2,  0, -1, -1,  0, -2, -1, -1, -1, -1, -2, -1,  1,  0, -3, -2,  0},
    /*R*/ { -1,  5,  0, -2, -3,  1,  0, -2,  0, -3, -2,  2, -1, -3, -2, -1, -1, -3, -2, -3},
    /*N*/ { -2,  0,  6,  1, -3,  0,  0,  0,  1, -3, -3,  0, -2, -3, -2,  1,  0, -4, -2, -3},
    /*D*/ { -2, -2,  1,  6, -3,  0,  2, -1, -1, -3, -4, -1, -3, -3, -1,  0, -1, -4, -3, -3},
    /*C*/ {  0, -3, -3, -3,  9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1},
    /*Q*/ { -1,  1,  0,  0, -3,  5,  2, -2,  0, -3, -2,  1,  0, -3, -1,  0, -1, -2, -1, -2},
    /*E*/ { -1,  0,  0,  2, -4,  2,  5, -2,  0, -3, -3,  1, -2, -3, -1,  0, -1, -3, -2, -2},
    /*G*/ {  0, -2,  0, -1, -3, -2, -2,  6, -2, -4, -4, -2, -3, -3, -2,  0, -2, -2, -3, -3},
    /*H*/ { -2,  0,  1, -1, -3,  0,  0, -2,  8, -3, -3, -1, -2, -1, -2, -1, -2, -2,  2, -3},
    /*I*/ { -1, -3, -3, -3, -1, -3, -3, -4, -3,  4,  2, -3,  1,  0, -3, -2, -1, -3, -1,  3},
    /*L*/ { -1, -2, -3, -4, -1, -2, -3, -4, -3,  2,  4, -2,  2,  0, -3, -2, -1, -2, -1,  1},
    /*K*/ { -1,  2,  0, -1, -3,  1,  1, -2, -1, -3, -2,  5, -1, -3, -1,  0, -1, -3, -2, -2},
    /*M*/ { -1, -1, -2, -3, -1,  0, -2, -3, -2,  1,  2, -1,  5,  0, -2, -1, -1, -1, -1,  1},
    /*F*/ { -2, -3, -3, -3, -2, -3, -3, -3, -1,  0,  0, -3,  0,  6, -4, -2, -2,  1,  3, -1},
    /*P*/ { -1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4,  7, -1, -1, -4, -3, -2},
    /*S*/ {  1, -1,  1,  0, -1,  0,  0,  0, -1, -2, -2,  0, -1, -2, -1,  4,  1, -3, -2, -2},
    /*T*/ {  0, -1,  0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1,  1,  5, -2, -2,  0},
    /*W*/ { -3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1,  1, -4, -3, -2, 11,  2, -3},
    /*Y*/ { -2, -2, -2, -3, -2, -1, -2, -3,  2, -1, -1, -2, -1,  3, -3, -2, -2,  2,  7, -1},
    /*V*/ {  0, -3, -3, -3, -1, -2, -2, -3, -3,  3,  1, -2,  1, -1, -2, -2,  0, -3, -1,  4},
};
/* Ambiguity rows over the column order ARNDCQEGHILKMFPSTWYV B Z X * U O J. */
static const char kAmbCols[] = "ARNDCQEGHILKMFPSTWYVBZX*UOJ";
static const char kAmbRows[] = "BZJ";
static const signed char kAmb[3][27] = {
    /*B*/ {-2, -1,  4,  4, -3,  0,  1, -1,  0, -3, -4,  0, -3, -3, -2,  0, -1, -4, -3, -3,  4,  0, -1, -4, -1, -1, -3},
    /*Z*/ {-1,  0,  0,  1, -3,  4,  4, -2,  0, -3, -3,  1, -1, -3, -1,  0, -1, -2, -2, -2,  0,  4, -1, -4, -1, -1, -3},
    /*J*/ {-1, -2, -3, -3, -1, -2, -3, -4, -3,  3,  3, -3,  2,  0, -3, -2, -1, -2, -1,  2, -3, -3, -1, -4, -1, -1,  3},
};
static const char kStdaa[] = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";

static int8_t g_blosum[OR_ALPHABET * OR_ALPHABET];
static pthread_once_t g_blosum_once = PTHREAD_ONCE_INIT;

static void blosum_init(void)
{
    int i, j;
    for (i = 0; i < OR_ALPHABET; ++i)
        for (j = 0; j < OR_ALPHABET; ++j) {
            char a = kStdaa[i], b = kStdaa[j];
            int v;
            if (a == '-' || b == '-')
                v = -128;
            else if (a == '*' || b == '*')
                v = (a == b) ? 1 : -4;
            else if (strchr("XUO", a) || strchr("XUO", b))
                v = -1;
            else if (strchr(kAmbRows, a))
                v = kAmb[strchr(kAmbRows, a) - kAmbRows][strchr(kAmbCols, b) - kAmbCols];
            else if (strchr(kAmbRows, b))
                v = kAmb[strchr(kAmbRows, b) - kAmbRows][strchr(kAmbCols, a) - kAmbCols];
            else
                v = kCore20[strchr(kCore, a) - kCore][strchr(kCore, b) - kCore];
            g_blosum[i * OR_ALPHABET + j] = (int8_t)v;
        }
}

const int8_t *or_blosum62(void)
{
    pthread_once(&g_blosum_once, blosum_init);
    return g_blosum;
}

/* ------------------------------------------------------------------------ */
int32_t or_diag_length(int32_t qlen, int32_t window)
{
    int32_t n = 1;
    while (n < qlen + window)
        n <<= 1;
    return n;
}

static uint32_t word_index(const uint8_t *w, int word_length)
{
    uint32_t idx = 0;
    for (int i = 0; i < word_length; ++i)
        idx = (idx << OR_CHARSIZE) | w[i];
    return idx;
}

/* Enumerates every word over the 28-letter alphabet and tests it against the
 * query word at position p: the neighbourhood definition itself, with no
 * pruning.  The exact query word is always a member. */
int or_lookup_build(const uint8_t *query, int32_t qlen, const int8_t *M,
                    int32_t threshold, int32_t word_length, or_lookup *out)
{
    memset(out, 0, sizeof(*out));
    if (word_length < 2 || word_length > 3 || qlen < 0 || qlen > 65535)
        return -1;
    const uint32_t nwords = 1u << (OR_CHARSIZE * word_length);
    uint32_t *count = (uint32_t *)calloc(nwords + 1, sizeof(uint32_t));
    if (!count)
        return -1;
    const int npos = qlen - word_length + 1;
    uint8_t w[3];
    for (int pass = 0; pass < 2; ++pass) {
        for (int p = 0; p < npos; ++p) {
            const uint8_t *qw = query + p;
            const uint32_t self = word_index(qw, word_length);
            const int8_t *r0 = M + qw[0] * OR_ALPHABET;
            const int8_t *r1 = M + qw[1] * OR_ALPHABET;
            const int8_t *r2 = word_length == 3 ? M + qw[2] * OR_ALPHABET : NULL;
            for (int a = 0; a < OR_ALPHABET; ++a)
                for (int b = 0; b < OR_ALPHABET; ++b)
                    for (int c = 0; c < (word_length == 3 ? OR_ALPHABET : 1); ++c) {
                        int score = r0[a] + r1[b] + (r2 ? r2[c] : 0);
                        w[0] = (uint8_t)a;
                        w[1] = (uint8_t)b;
                        w[2] = (uint8_t)c;
                        uint32_t idx = word_index(w, word_length);
                        if (score >= threshold || idx == self) {
                            if (pass == 0)
                                count[idx]++;
                            else
                                out->entries[out->cell_start[idx] + count[idx]++] = (uint16_t)p;
                        }
                    }
        }
        if (pass == 0) {
            out->cell_start = (uint32_t *)malloc((nwords + 1) * sizeof(uint32_t));
            if (!out->cell_start) {
                free(count);
                return -1;
            }
            uint32_t acc = 0, mx = 0;
            for (uint32_t i = 0; i < nwords; ++i) {
                out->cell_start[i] = acc;
                acc += count[i];
                if (count[i] > mx)
                    mx = count[i];
                count[i] = 0;
            }
            out->cell_start[nwords] = acc;
            out->entries = (uint16_t *)malloc((acc ? acc : 1) * sizeof(uint16_t));
            if (!out->entries) {
                free(count);
                free(out->cell_start);
                out->cell_start = NULL;
                return -1;
            }
            out->max_cell = mx;
        }
    }
    free(count);
    out->word_length = word_length;
    out->nwords = nwords;
    return 0;
}

void or_lookup_free(or_lookup *lk)
{
    free(lk->cell_start);
    free(lk->entries);
    memset(lk, 0, sizeof(*lk));
}

/* ------------------------------------------------------------------------ */
/* Left X-drop extension from (s_off, q_off) towards the sequence starts. */
or_ext_left or_extend_left(const int8_t *M, const uint8_t *subject,
                           const uint8_t *query, int32_t s_off, int32_t q_off,
                           int32_t dropoff, int32_t maxscore)
{
    int32_t n = OR_MIN(s_off, q_off);
    int32_t best_i = n + 1;
    int32_t score = maxscore;
    for (int32_t i = n; i >= 0; --i) {
        score += M[query[q_off - n + i] * OR_ALPHABET + subject[s_off - n + i]];
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (maxscore - score >= dropoff)
            break;
    }
    or_ext_left r = {maxscore, n - best_i + 1};
    return r;
}

/* Right X-drop extension; also stops once the running score is <= 0. */
or_ext_right or_extend_right(const int8_t *M, const uint8_t *subject,
                             const uint8_t *query, int32_t s_off, int32_t q_off,
                             int32_t dropoff, int32_t maxscore,
                             int32_t subject_length, int32_t query_length)
{
    int32_t n = OR_MIN(subject_length - s_off, query_length - q_off);
    int32_t best_i = -1;
    int32_t score = maxscore;
    int32_t i;
    for (i = 0; i < n; ++i) {
        score += M[query[q_off + i] * OR_ALPHABET + subject[s_off + i]];
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (score <= 0 || maxscore - score >= dropoff)
            break;
    }
    or_ext_right r = {maxscore, best_i + 1, s_off + i};
    return r;
}

or_ext_two or_extend_two_hit(const int8_t *M, const uint8_t *subject,
                             int32_t subject_length, const uint8_t *query,
                             int32_t query_length, int32_t s_left_off,
                             int32_t s_right_off, int32_t q_right_off,
                             int32_t dropoff, int32_t word_size)
{
    int32_t left_score = 0, right_score = 0, right_d = 0, left_d;
    int32_t score = 0;
    /* best prefix of the second word: right anchor is one past its end */
    for (int32_t i = 0; i < word_size; ++i) {
        score += M[query[q_right_off + i] * OR_ALPHABET + subject[s_right_off + i]];
        if (score > left_score) {
            left_score = score;
            right_d = i + 1;
        }
    }
    q_right_off += right_d;
    s_right_off += right_d;
    right_d = 0;

    or_ext_two r;
    r.right_extend = 0;
    r.s_last_off = s_right_off;

    or_ext_left L = or_extend_left(M, subject, query, s_right_off - 1,
                                   q_right_off - 1, dropoff, 0);
    left_score = L.maxscore;
    left_d = L.length;

    if (left_d >= s_right_off - s_left_off) {
        r.right_extend = 1;
        or_ext_right R = or_extend_right(M, subject, query, s_right_off,
                                         q_right_off, dropoff, left_score,
                                         subject_length, query_length);
        right_score = R.maxscore;
        right_d = R.length;
        r.s_last_off = R.s_last_off;
    }
    r.hsp_q = q_right_off - left_d;
    r.hsp_s = s_right_off - left_d;
    r.hsp_len = left_d + right_d;
    r.score = OR_MAX(left_score, right_score);
    return r;
}

/* ------------------------------------------------------------------------ */
static void diag_clear(or_diag *diag, int32_t len, int32_t window)
{
    for (int32_t i = 0; i < len; ++i) {
        diag[i].flag = 0;
        diag[i].last_hit = -window;
    }
}

int32_t or_scan_subject(const or_lookup *lk, const uint8_t *query, int32_t qlen,
                        const int8_t *M, const or_params *p,
                        const uint8_t *subject, int32_t slen,
                        or_diag *diag, int32_t *diag_offset_io,
                        or_hsp *out, int32_t cap, int32_t counters[3])
{
    const int32_t word = p->word_length;
    const int32_t window = p->window;
    const int32_t mask = p->diag_mask;
    const int32_t diag_offset = *diag_offset_io;
    int32_t hits = 0, extended = 0, saved = 0;

    for (int32_t s = 0; s <= slen - word; ++s) {
        const uint32_t idx = word_index(subject + s, word);
        const uint32_t b = lk->cell_start[idx], e = lk->cell_start[idx + 1];
        hits += (int32_t)(e - b);
        for (uint32_t k = b; k < e; ++k) {
            const int32_t q = lk->entries[k];
            const int32_t d = (q - s) & mask;
            or_diag *st = &diag[d];
            if (st->flag) {
                if (s + diag_offset < st->last_hit)
                    continue;
                st->last_hit = s + diag_offset;
                st->flag = 0;
                continue;
            }
            const int32_t last = st->last_hit - diag_offset;
            const int32_t diff = s - last;
            if (diff >= window) {
                st->last_hit = s + diag_offset;
                continue;
            }
            if (diff < word)
                continue;
            or_ext_two x = or_extend_two_hit(M, subject, slen, query, qlen,
                                             last + word, s, q, p->x_dropoff, word);
            ++extended;
            if (x.score >= p->cutoff) {
                if (saved < cap) {
                    or_hsp h = {x.hsp_q, x.hsp_s, q, s, x.hsp_len, x.score};
                    out[saved] = h;
                }
                ++saved;
            }
            if (x.right_extend) {
                st->flag = 1;
                st->last_hit = x.s_last_off - (word - 1) + diag_offset;
            } else {
                st->last_hit = s + diag_offset;
            }
        }
    }
    /* Blast_ExtendWordExit: advance the offset, clearing the table when the
     * 31-bit last_hit field would run out. */
    if (diag_offset >= INT32_MAX / 4) {
        *diag_offset_io = window;
        diag_clear(diag, mask + 1, window);
    } else {
        *diag_offset_io = diag_offset + slen + window;
    }
    counters[0] = hits;
    counters[1] = saved;
    counters[2] = extended;
    return saved;
}

/* ------------------------------------------------------------------------ */
typedef struct run_shared {
    const or_lookup *lk;
    const uint8_t *query;
    int32_t qlen;
    const int8_t *M;
    const or_params *p;
    const uint8_t *residues;
    const uint64_t *off;
    const uint32_t *len;
    uint32_t n;
    int32_t *counters;
    atomic_uint next_chunk;
    /* per-subject record lists, stitched into CSR afterwards */
    or_hsp **lists;
    int failed;
} run_shared;

#define OR_CHUNK 4096u

static void *run_worker(void *arg)
{
    run_shared *sh = (run_shared *)arg;
    const int32_t dlen = sh->p->diag_mask + 1;
    or_diag *diag = (or_diag *)malloc((size_t)dlen * sizeof(or_diag));
    int32_t cap = 64;
    or_hsp *buf = (or_hsp *)malloc((size_t)cap * sizeof(or_hsp));
    if (!diag || !buf) {
        sh->failed = 1;
        free(diag);
        free(buf);
        return NULL;
    }
    int32_t diag_offset = sh->p->window;
    diag_clear(diag, dlen, sh->p->window);
    for (;;) {
        uint32_t c = atomic_fetch_add(&sh->next_chunk, 1u);
        uint64_t first = (uint64_t)c * OR_CHUNK;
        if (first >= sh->n)
            break;
        uint32_t last = (uint32_t)OR_MIN((uint64_t)sh->n, first + OR_CHUNK);
        for (uint32_t i = (uint32_t)first; i < last; ++i) {
            int32_t *cnt = sh->counters + 3 * (size_t)i;
            const uint8_t *subj = sh->residues + sh->off[i];
            int32_t saved;
            for (;;) {
                saved = or_scan_subject(sh->lk, sh->query, sh->qlen, sh->M, sh->p,
                                        subj, (int32_t)sh->len[i], diag,
                                        &diag_offset, buf, cap, cnt);
                if (saved <= cap)
                    break;
                /* more records than the buffer holds: grow it and rescan the
                 * subject from a cleared table (a subject's results do not
                 * depend on the table state left by earlier subjects) */
                cap = saved * 2;
                free(buf);
                buf = (or_hsp *)malloc((size_t)cap * sizeof(or_hsp));
                if (!buf) {
                    sh->failed = 1;
                    free(diag);
                    return NULL;
                }
                diag_offset = sh->p->window;
                diag_clear(diag, dlen, sh->p->window);
            }
            if (saved > 0) {
                or_hsp *l = (or_hsp *)malloc((size_t)saved * sizeof(or_hsp));
                if (!l) {
                    sh->failed = 1;
                    break;
                }
                memcpy(l, buf, (size_t)saved * sizeof(or_hsp));
                sh->lists[i] = l;
            }
        }
    }
    free(diag);
    free(buf);
    return NULL;
}

int or_run_db(const or_lookup *lk, const uint8_t *query, int32_t qlen,
              const int8_t *M, const or_params *p,
              const uint8_t *residues, const uint64_t *off, const uint32_t *len,
              uint32_t n, int nthreads, int32_t *counters,
              uint64_t *rec_offsets, or_hsp **records)
{
    if (nthreads <= 0)
        nthreads = 1;
    run_shared sh;
    memset(&sh, 0, sizeof(sh));
    sh.lk = lk;
    sh.query = query;
    sh.qlen = qlen;
    sh.M = M;
    sh.p = p;
    sh.residues = residues;
    sh.off = off;
    sh.len = len;
    sh.n = n;
    sh.counters = counters;
    atomic_init(&sh.next_chunk, 0u);
    sh.lists = (or_hsp **)calloc(n ? n : 1, sizeof(or_hsp *));
    if (!sh.lists)
        return -1;
    pthread_t *th = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    int started = 0;
    for (int t = 0; t < nthreads; ++t)
        if (pthread_create(&th[t], NULL, run_worker, &sh) == 0)
            ++started;
    if (started == 0)
        run_worker(&sh);
    for (int t = 0; t < started; ++t)
        pthread_join(th[t], NULL);
    free(th);

    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        rec_offsets[i] = total;
        total += (uint64_t)counters[3 * (size_t)i + 1];
    }
    rec_offsets[n] = total;
    or_hsp *rec = (or_hsp *)malloc((total ? total : 1) * sizeof(or_hsp));
    if (!rec || sh.failed) {
        for (uint32_t i = 0; i < n; ++i)
            free(sh.lists[i]);
        free(sh.lists);
        free(rec);
        return -1;
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t c = rec_offsets[i + 1] - rec_offsets[i];
        if (c)
            memcpy(rec + rec_offsets[i], sh.lists[i], c * sizeof(or_hsp));
        free(sh.lists[i]);
    }
    free(sh.lists);
    *records = rec;
    return 0;
}

void or_free(void *ptr) { free(ptr); }
