"""A second, deliberately naive restatement of the word finder, used only to
check the C oracle on tiny inputs.

It shares no code or data structure with oracle/oracle.c: words are matched by
scoring every (query position, subject position) pair directly instead of
through a lookup table, the diagonal state is a dict in plain subject
coordinates (no running offset), and the extensions are written as while
loops over explicit coordinates.  Semantics follow gpu_blastp_kernel.cu:
scan/two-hit :151-301, s_BlastAaExtendTwoHit :341-415, s_BlastAaExtendLeft
:435-481, s_BlastAaExtendRight :503-549.
"""


def word_hits(query, subject, M, word=3, T=11):
    """All (s, q) word pairs, s-major then q ascending, scoring >= T or equal."""
    out = []
    for s in range(len(subject) - word + 1):
        sw = tuple(int(c) for c in subject[s:s + word])
        for q in range(len(query) - word + 1):
            qw = tuple(int(c) for c in query[q:q + word])
            sc = sum(int(M[qw[i]][sw[i]]) for i in range(word))
            if sc >= T or qw == sw:
                out.append((s, q))
    return out


def two_hit_extension(query, subject, s_left, s_right, q_right, xdrop, word, M):
    sc = best = rd = 0
    for i in range(word):
        sc += int(M[query[q_right + i]][subject[s_right + i]])
        if sc > best:
            best, rd = sc, i + 1
    qr, sr = q_right + rd, s_right + rd
    run = mx = L = 0
    k = 0
    while sr - 1 - k >= 0 and qr - 1 - k >= 0:
        run += int(M[query[qr - 1 - k]][subject[sr - 1 - k]])
        if run > mx:
            mx, L = run, k + 1
        if mx - run >= xdrop:
            break
        k += 1
    right = L >= sr - s_left
    R, total, s_last = 0, mx, sr
    if right:
        run = mx2 = mx
        best_k, k = -1, 0
        while sr + k < len(subject) and qr + k < len(query):
            run += int(M[query[qr + k]][subject[sr + k]])
            if run > mx2:
                mx2, best_k = run, k
            if run <= 0 or mx2 - run >= xdrop:
                break
            k += 1
        R, total, s_last = best_k + 1, mx2, sr + k
    return qr - L, sr - L, L + R, total, right, s_last


def scan(query, subject, M, word=3, T=11, window=40, xdrop=16, cutoff=41, diag_len=None):
    """Returns (hits, saved, extended, records) for one subject, records as
    (hsp_q, hsp_s, q_off, s_off, len, score) in emission order."""
    if diag_len is None:
        diag_len = 1
        while diag_len < len(query) + window:
            diag_len *= 2
    mask = diag_len - 1
    state = {}
    hits = extended = 0
    recs = []
    for s, q in word_hits(query, subject, M, word, T):
        hits += 1
        d = (q - s) & mask
        last, flag = state.get(d, (-window, False))
        if flag:
            if s >= last:
                state[d] = (s, False)
            continue
        diff = s - last
        if diff >= window:
            state[d] = (s, False)
            continue
        if diff < word:
            continue
        hq, hs, ln, score, right, s_last = two_hit_extension(query, subject, last + word, s, q, xdrop, word, M)
        extended += 1
        if score >= cutoff:
            recs.append((hq, hs, q, s, ln, score))
        state[d] = (s_last - (word - 1), True) if right else (s, False)
    return hits, len(recs), extended, recs
