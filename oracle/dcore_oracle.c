/*
 * dcore_oracle.c -- CPU oracle for D-core ((k,l)-core) decomposition.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Semantics (restated from the reference, llysdal/d-core-cuda):
 *   - One peel pass for a fixed k (scan main.cu:27-62, process main.cu:72-141,
 *     level loop KListInplace main.cu:171-195): start at level L = 0 with all
 *     vertices alive.  At level L repeatedly remove every alive v with
 *     out-degree(alive) <= L or in-degree(alive) < k, giving it l[v] = L, until
 *     none is left; then L := L + 1; stop when nothing is alive.  The return
 *     value is the last level (KListInplace returns level - 1).
 *     The reference tests out_deg == L (main.cu:44); that is the same test
 *     because every alive vertex has out-degree >= L when level L starts.
 *   - kmax pass: the same peel with in and out swapped and k = 0
 *     (swapInOut, src/common.h:114-120; dcore main.cu:218-220).
 *   - Rows k = 0..kmax (dcore main.cu:238-244), lmax = max over rows.
 *
 * Implementation: per pass a bucket queue keyed by the current out-degree
 * with lazy entries (a vertex is re-filed into a lower bucket each time its
 * key drops above the current level), so one pass costs O(V + E + maxdeg)
 * instead of the reference's O(levels * V) scans.
 *
 * Pinning (each is a test under tests/):
 *   dco_klist / dco_kmax / dco_lrows
 *     - the reference's golden vector dataset/digraph(.decomp)
 *       (tests/test_oracle.py::test_fixture_decomp);
 *     - brute force from the definition (iterative removal for every (k,l))
 *       on random and exhaustive tiny digraphs;
 *     - textbook reductions: kmaxes = in-degree core numbers and
 *       lmaxes[0] = out-degree core numbers by an independent
 *       Batagelj-Zaversnik implementation, and on symmetric digraphs both
 *       equal networkx.core_number of the undirected graph;
 *     - closed forms (complete digraph, directed cycle, DAGs, stars,
 *       bidirected K_{a,b}, circulants);
 *     - reverse duality lmaxes_G[k][v] = max{l : v in D_{l,k}(G^R)};
 *     - invariants (monotone in k, zero above kmaxes[v], degree bounds,
 *       core property and maximality of every level set).
 *   dco_klist_hindex / dco_kmax_hindex
 *     - equality with the peel on every test graph (two different
 *       algorithms, fixpoint argument in the A.3 design, main.cu:982-1242).
 *   dco_state_hash (rows peeled on their (k,0)-cores, streamed into a hash)
 *     - equality, part by part, with the hash formula applied to the dense
 *       output of dco_kmax / dco_lrows on the fixture, random graphs and
 *       closed forms (tests/test_oracle.py::test_streaming_hash_*); a
 *       changed cell moves the hash.
 */
#include "dcore_oracle.h"

#include <stdlib.h>
#include <string.h>

/* DCO_PLANT (tests only): compiles a deliberately wrong variant so the
 * tests can show that their checks catch plausible mistakes.
 *   1: constraint threshold off by one;  2: key decrements sent to the
 *   wrong side;  3: levels assigned one too high.  Unset in every build
 *   that is used as the checker. */
#ifndef DCO_PLANT
#define DCO_PLANT 0
#endif
#if DCO_PLANT == 1
#define CFAIL(x, k) ((x) <= (k))
#else
#define CFAIL(x, k) ((x) < (k))
#endif
#if DCO_PLANT == 3
#define LEVEL(L) ((L) + 1)
#else
#define LEVEL(L) (L)
#endif
#ifdef _OPENMP
#include <omp.h>
#endif

/* Generic peel.  Key lists K(v) decide the level (out-lists for an l-pass);
 * constraint lists C(v) decide membership with threshold k (in-lists for an
 * l-pass).  Removing v lowers the key degree of every w in C(v) (v was in
 * K(w)) and the constraint degree of every u in K(v). */
/* Entries of a sorted list below n (its prefix of members). */
static uint64_t count_below(const uint32_t* a, uint64_t len, uint32_t n) {
    uint64_t lo = 0, hi = len;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < n)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* prefix == 0: the graph is vertices [0, V) with whole lists.
 * prefix != 0: only vertices [0, V) of a larger graph whose lists are
 * sorted take part (the (k,0)-core in kmax-descending rank order); each
 * list is read up to its first entry >= V. */
static int32_t peel_n(uint32_t V, const uint64_t* koff, const uint32_t* knbr,
                      const uint64_t* coff, const uint32_t* cnbr, int32_t k,
                      int32_t* l, int prefix) {
    if (V == 0) return -1;
    int32_t* kd = (int32_t*)malloc(sizeof(int32_t) * V);
    int32_t* cd = (int32_t*)malloc(sizeof(int32_t) * V);
    uint8_t* state = (uint8_t*)calloc(V, 1); /* 0 alive, 1 removed */
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * V);
    int32_t maxkd = 0;
    uint64_t csum = 0;
    for (uint32_t v = 0; v < V; ++v) {
        if (prefix) {
            kd[v] = (int32_t)count_below(knbr + koff[v], koff[v + 1] - koff[v], V);
            cd[v] = (int32_t)count_below(cnbr + coff[v], coff[v + 1] - coff[v], V);
        } else {
            kd[v] = (int32_t)(koff[v + 1] - koff[v]);
            cd[v] = (int32_t)(coff[v + 1] - coff[v]);
        }
        csum += (uint64_t)cd[v];
        if (kd[v] > maxkd) maxkd = kd[v];
    }
    /* lazy buckets: at most V initial entries plus one per key decrement */
    uint64_t cap = (uint64_t)V + csum;
    uint32_t* pv = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint64_t* pnext = (uint64_t*)malloc(sizeof(uint64_t) * cap);
    uint64_t* head = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)maxkd + 1));
    const uint64_t NIL = ~(uint64_t)0;
    for (int32_t d = 0; d <= maxkd; ++d) head[d] = NIL;
    uint64_t np = 0;
    for (uint32_t v = 0; v < V; ++v) {
        pv[np] = v;
        pnext[np] = head[kd[v]];
        head[kd[v]] = np++;
    }
    uint32_t removed = 0, sp = 0;
    int32_t L = 0, last = -1;
    /* level 0 also takes every vertex failing the constraint */
    for (uint32_t v = 0; v < V; ++v)
        if (CFAIL(cd[v], k)) {
            state[v] = 1;
            l[v] = LEVEL(0);
            stack[sp++] = v;
        }
    for (;; ++L) {
        if (L <= maxkd) {
            for (uint64_t e = head[L]; e != NIL; e = pnext[e]) {
                uint32_t v = pv[e];
                if (state[v] == 0 && kd[v] <= L) {
                    state[v] = 1;
                    l[v] = LEVEL(L);
                    stack[sp++] = v;
                }
            }
        }
        while (sp > 0) {
            uint32_t v = stack[--sp];
            ++removed;
            last = L;
#if DCO_PLANT == 2
            for (uint64_t o = koff[v]; o < koff[v + 1]; ++o) {
                uint32_t w = knbr[o];
#else
            for (uint64_t o = coff[v]; o < coff[v + 1]; ++o) {
                uint32_t w = cnbr[o];
#endif
                if (w >= V) break; /* prefix mode: past the members */
                if (state[w]) continue;
                if (--kd[w] <= L) {
                    state[w] = 1;
                    l[w] = LEVEL(L);
                    stack[sp++] = w;
                } else {
                    pv[np] = w;
                    pnext[np] = head[kd[w]];
                    head[kd[w]] = np++;
                }
            }
            for (uint64_t o = koff[v]; o < koff[v + 1]; ++o) {
                uint32_t u = knbr[o];
                if (u >= V) break; /* prefix mode: past the members */
                if (state[u]) continue;
                if (CFAIL(--cd[u], k)) {
                    state[u] = 1;
                    l[u] = LEVEL(L);
                    stack[sp++] = u;
                }
            }
        }
        if (removed == V) break;
    }
    free(kd);
    free(cd);
    free(state);
    free(stack);
    free(pv);
    free(pnext);
    free(head);
    return last;
}

static int32_t peel(uint32_t V, const uint64_t* koff, const uint32_t* knbr,
                    const uint64_t* coff, const uint32_t* cnbr, int32_t k,
                    int32_t* l) {
    return peel_n(V, koff, knbr, coff, cnbr, k, l, 0);
}

int32_t dco_klist(uint32_t V, const uint64_t* out_off, const uint32_t* out_nbr,
                  const uint64_t* in_off, const uint32_t* in_nbr, int32_t k,
                  int32_t* l_out) {
    return peel(V, out_off, out_nbr, in_off, in_nbr, k, l_out);
}

int32_t dco_kmax(uint32_t V, const uint64_t* out_off, const uint32_t* out_nbr,
                 const uint64_t* in_off, const uint32_t* in_nbr,
                 int32_t* kmaxes) {
    return peel(V, in_off, in_nbr, out_off, out_nbr, 0, kmaxes);
}

int32_t dco_lrows(uint32_t V, const uint64_t* out_off, const uint32_t* out_nbr,
                  const uint64_t* in_off, const uint32_t* in_nbr, int32_t kmax,
                  int32_t* lmaxes, int nthreads) {
    int32_t lmax = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(max : lmax) num_threads(nthreads)
#endif
    for (int32_t k = 0; k <= kmax; ++k) {
        int32_t lk = peel(V, out_off, out_nbr, in_off, in_nbr, k,
                          lmaxes + (size_t)k * V);
        if (lk > lmax) lmax = lk;
    }
    (void)nthreads;
    return lmax;
}

/* h-index of the multiset {min(cap, x_i)}: the largest h with at least h
 * entries >= h.  hist must hold cap+1 zeroed counters; it is re-zeroed. */
static int32_t hindex_capped(const int32_t* vals, const uint32_t* ids,
                             uint64_t n, int32_t cap, int32_t* hist,
                             const uint8_t* member, int32_t need) {
    /* need == 0: h-index; need > 0: largest h with >= need entries >= h */
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t w = ids[i];
        if (member && !member[w]) continue;
        int32_t x = vals[w];
        hist[x < cap ? x : cap]++;
    }
    int32_t cnt = 0, h = 0;
    for (int32_t i = cap; i >= 0; --i) {
        cnt += hist[i];
        if (cnt >= (need ? need : i)) {
            h = i;
            break;
        }
    }
    for (int32_t i = 0; i <= cap; ++i) hist[i] = 0;
    return h;
}

int32_t dco_klist_hindex(uint32_t V, const uint64_t* out_off,
                         const uint32_t* out_nbr, const uint64_t* in_off,
                         const uint32_t* in_nbr, const int32_t* kmaxes,
                         int32_t k, int32_t* l_out, int32_t* iters) {
    uint8_t* member = (uint8_t*)malloc(V ? V : 1);
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (V ? V : 1));
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * (V ? V : 1));
    int32_t maxd = 0;
    for (uint32_t v = 0; v < V; ++v) member[v] = kmaxes[v] >= k;
    for (uint32_t v = 0; v < V; ++v) {
        int32_t d = 0;
        if (member[v])
            for (uint64_t o = out_off[v]; o < out_off[v + 1]; ++o)
                d += member[out_nbr[o]];
        cur[v] = d;
        if (d > maxd) maxd = d;
    }
    int32_t* hist = (int32_t*)calloc((size_t)maxd + 1, sizeof(int32_t));
    int32_t it = 0;
    for (;;) {
        int changed = 0;
        ++it;
        for (uint32_t v = 0; v < V; ++v) {
            if (!member[v]) {
                nxt[v] = 0;
                continue;
            }
            int32_t lv = cur[v];
            int32_t ho = hindex_capped(cur, out_nbr + out_off[v],
                                       out_off[v + 1] - out_off[v], lv, hist,
                                       member, 0);
            int32_t hi = k == 0 ? lv
                                : hindex_capped(cur, in_nbr + in_off[v],
                                                in_off[v + 1] - in_off[v], lv,
                                                hist, member, k);
            int32_t nv = lv;
            if (ho < nv) nv = ho;
            if (hi < nv) nv = hi;
            nxt[v] = nv;
            changed |= nv != lv;
        }
        int32_t* t = cur;
        cur = nxt;
        nxt = t;
        if (!changed) break;
    }
    int32_t lmax = 0;
    for (uint32_t v = 0; v < V; ++v) {
        l_out[v] = cur[v];
        if (cur[v] > lmax) lmax = cur[v];
    }
    if (iters) *iters = it;
    free(member);
    free(cur);
    free(nxt);
    free(hist);
    return lmax;
}

int32_t dco_kmax_hindex(uint32_t V, const uint64_t* in_off,
                        const uint32_t* in_nbr, int32_t* kmaxes,
                        int32_t* iters) {
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * (V ? V : 1));
    int32_t maxd = 0;
    for (uint32_t v = 0; v < V; ++v) {
        kmaxes[v] = (int32_t)(in_off[v + 1] - in_off[v]);
        if (kmaxes[v] > maxd) maxd = kmaxes[v];
    }
    int32_t* hist = (int32_t*)calloc((size_t)maxd + 1, sizeof(int32_t));
    int32_t it = 0;
    for (;;) {
        int changed = 0;
        ++it;
        for (uint32_t v = 0; v < V; ++v) {
            int32_t kv = kmaxes[v];
            int32_t h = hindex_capped(kmaxes, in_nbr + in_off[v],
                                      in_off[v + 1] - in_off[v], kv, hist,
                                      NULL, 0);
            nxt[v] = h < kv ? h : kv;
            changed |= nxt[v] != kv;
        }
        memcpy(kmaxes, nxt, sizeof(int32_t) * V);
        if (!changed) break;
    }
    int32_t kmax = V ? 0 : -1;
    for (uint32_t v = 0; v < V; ++v)
        if (kmaxes[v] > kmax) kmax = kmaxes[v];
    if (iters) *iters = it;
    free(nxt);
    free(hist);
    return kmax;
}

/* ---- streaming state hash ---------------------------------------------
 * The hash that dcx_state_hash defines (include/dcx.h): the 64-bit sum over
 * every vertex v and k = -1..kmax of mix64(v*G1 + (k+1)*G2 + value), value
 * = kmaxes[v] for k = -1 and lmaxes[k][v] otherwise (0 outside the
 * (k,0)-core), mix64 the SplitMix64 finaliser.  Each row is computed, added
 * and dropped, so a scale-24 decomposition needs O(V) per thread instead of
 * (kmax+1) * V cells.  Row k is peeled on its (k,0)-core alone: vertices
 * relabelled by kmax descending (ties by id) make the core the rank prefix
 * [0, V_k), and lists sorted by rank make each member's member-neighbours
 * a list prefix (peel_n in prefix mode).  The values are those of the
 * whole-graph peel: a vertex outside the (k,0)-core fails the constraint at
 * level 0 there and cannot affect a member's degrees afterwards. */
static uint64_t hmix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
#define HG1 0x9E3779B97F4A7C15ULL
#define HG2 0xC2B2AE3D27D4EB4FULL

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

/* Lists of one side in rank space, each sorted. */
static void rank_lists(uint32_t V, const uint64_t* off, const uint32_t* nbr,
                       const uint32_t* order, const uint32_t* rank,
                       uint64_t* roff, uint32_t* rnbr) {
    roff[0] = 0;
    for (uint32_t r = 0; r < V; ++r) {
        uint32_t v = order[r];
        roff[r + 1] = roff[r] + (off[v + 1] - off[v]);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096)
#endif
    for (int64_t r = 0; r < (int64_t)V; ++r) {
        uint32_t v = order[r];
        uint64_t o = roff[r];
        for (uint64_t i = off[v]; i < off[v + 1]; ++i) rnbr[o++] = rank[nbr[i]];
        qsort(rnbr + roff[r], roff[r + 1] - roff[r], sizeof(uint32_t), cmp_u32);
    }
}

uint64_t dco_state_hash(uint32_t V, const uint64_t* out_off, const uint32_t* out_nbr,
                        const uint64_t* in_off, const uint32_t* in_nbr, int nthreads,
                        int32_t* kmax_out, int32_t* lmax_out, uint64_t* row_hash) {
    if (kmax_out) *kmax_out = -1;
    if (lmax_out) *lmax_out = -1;
    if (V == 0) return 0;
    int32_t* km = (int32_t*)malloc(sizeof(int32_t) * V);
    int32_t kmax = peel(V, in_off, in_nbr, out_off, out_nbr, 0, km);
    /* order by kmax descending, ties by id */
    uint32_t* cnt = (uint32_t*)calloc((size_t)kmax + 2, sizeof(uint32_t));
    for (uint32_t v = 0; v < V; ++v) cnt[kmax - km[v]]++;
    uint32_t* start = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)kmax + 2));
    uint32_t acc = 0;
    for (int32_t j = 0; j <= kmax; ++j) {
        start[j] = acc;
        acc += cnt[j];
    }
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * V);
    uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * V);
    for (uint32_t v = 0; v < V; ++v) {
        uint32_t r = start[kmax - km[v]]++;
        order[r] = v;
        rank[v] = r;
    }
    /* V_k = members of the (k,0)-core = ranks below vk[k] */
    uint32_t* vk = (uint32_t*)calloc((size_t)kmax + 2, sizeof(uint32_t));
    for (uint32_t v = 0; v < V; ++v) vk[km[v]]++;
    for (int32_t k = kmax - 1; k >= 0; --k) vk[k] += vk[k + 1];
    uint64_t* coff_o = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)V + 1));
    uint64_t* coff_i = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)V + 1));
    uint32_t* cnbr_o = (uint32_t*)malloc(sizeof(uint32_t) * (out_off[V] ? out_off[V] : 1));
    uint32_t* cnbr_i = (uint32_t*)malloc(sizeof(uint32_t) * (in_off[V] ? in_off[V] : 1));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    omp_set_num_threads(nthreads);
#endif
    rank_lists(V, out_off, out_nbr, order, rank, coff_o, cnbr_o);
    rank_lists(V, in_off, in_nbr, order, rank, coff_i, cnbr_i);
    uint64_t total = 0;
    for (uint32_t v = 0; v < V; ++v) total += hmix(v * HG1 + (uint64_t)(uint32_t)km[v]);
    if (row_hash) row_hash[0] = total;
    int32_t lmax = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) reduction(max : lmax) num_threads(nthreads)
#endif
    for (int32_t k = 0; k <= kmax; ++k) {
        uint32_t n = vk[k];
        int32_t* l = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
        int32_t lk = peel_n(n, coff_o, cnbr_o, coff_i, cnbr_i, k, l, 1);
        uint64_t h = 0;
        const uint64_t kk = (uint64_t)(k + 1) * HG2;
        for (uint32_t r = 0; r < V; ++r)
            h += hmix(order[r] * HG1 + kk + (r < n ? (uint64_t)(uint32_t)l[r] : 0));
        free(l);
        if (row_hash) row_hash[k + 1] = h;
        total += h;
        if (lk > lmax) lmax = lk;
    }
    if (kmax_out) *kmax_out = kmax;
    if (lmax_out) *lmax_out = lmax;
    free(km);
    free(cnt);
    free(start);
    free(order);
    free(rank);
    free(vk);
    free(coff_o);
    free(coff_i);
    free(cnbr_o);
    free(cnbr_i);
    return total;
}

/* ---- timed sample (the CPU baseline of a benchmark) --------------------
 * Runs, on the host threads, the parts of one oracle decomposition: the
 * kmax pass, the kmax-rank relabelling of dco_state_hash, and the rows
 * k = 0, stride, 2*stride, ... (each peeled on its (k,0)-core, rows in
 * parallel over nthreads); the rest of the rows are not run.  Returns kmax;
 * t[0..2] = wall seconds of the three parts, w[0] = sum of (8 E_k + 20 V_k)
 * over the sampled rows and w[1] over all rows (the algorithmic bytes of
 * the rows, the unit by which a caller scales the sample up to the whole
 * decomposition). */
#ifdef _OPENMP
static double now_s(void) { return omp_get_wtime(); }
#else
#include <time.h>
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
#endif

int32_t dco_time_sample(uint32_t V, const uint64_t* out_off, const uint32_t* out_nbr,
                        const uint64_t* in_off, const uint32_t* in_nbr, int nthreads,
                        int32_t stride, double* t, double* w) {
    t[0] = t[1] = t[2] = 0;
    w[0] = w[1] = 0;
    if (V == 0) return -1;
    if (stride < 1) stride = 1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    omp_set_num_threads(nthreads);
#endif
    double t0 = now_s();
    int32_t* km = (int32_t*)malloc(sizeof(int32_t) * V);
    int32_t kmax = peel(V, in_off, in_nbr, out_off, out_nbr, 0, km);
    double t1 = now_s();
    uint32_t* cnt = (uint32_t*)calloc((size_t)kmax + 2, sizeof(uint32_t));
    for (uint32_t v = 0; v < V; ++v) cnt[kmax - km[v]]++;
    uint32_t* start = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)kmax + 2));
    uint32_t acc = 0;
    for (int32_t j = 0; j <= kmax; ++j) {
        start[j] = acc;
        acc += cnt[j];
    }
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * V);
    uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * V);
    for (uint32_t v = 0; v < V; ++v) {
        uint32_t r = start[kmax - km[v]]++;
        order[r] = v;
        rank[v] = r;
    }
    uint32_t* vk = (uint32_t*)calloc((size_t)kmax + 2, sizeof(uint32_t));
    for (uint32_t v = 0; v < V; ++v) vk[km[v]]++;
    for (int32_t k = kmax - 1; k >= 0; --k) vk[k] += vk[k + 1];
    uint64_t* coff_o = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)V + 1));
    uint64_t* coff_i = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)V + 1));
    uint32_t* cnbr_o = (uint32_t*)malloc(sizeof(uint32_t) * (out_off[V] ? out_off[V] : 1));
    uint32_t* cnbr_i = (uint32_t*)malloc(sizeof(uint32_t) * (in_off[V] ? in_off[V] : 1));
    rank_lists(V, out_off, out_nbr, order, rank, coff_o, cnbr_o);
    rank_lists(V, in_off, in_nbr, order, rank, coff_i, cnbr_i);
    double t2 = now_s();
    /* E_k: edges whose endpoints both have kmax >= k */
    uint64_t* ek = (uint64_t*)calloc((size_t)kmax + 2, sizeof(uint64_t));
    for (uint32_t u = 0; u < V; ++u)
        for (uint64_t o = out_off[u]; o < out_off[u + 1]; ++o) {
            int32_t a = km[u], b = km[out_nbr[o]];
            ek[a < b ? a : b]++;
        }
    for (int32_t k = kmax - 1; k >= 0; --k) ek[k] += ek[k + 1];
    for (int32_t k = 0; k <= kmax; ++k) {
        double wk = 8.0 * (double)ek[k] + 20.0 * (double)vk[k];
        w[1] += wk;
        if (k % stride == 0) w[0] += wk;
    }
    double t3 = now_s();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int32_t k = 0; k <= kmax; k += stride) {
        uint32_t n = vk[k];
        int32_t* l = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
        peel_n(n, coff_o, cnbr_o, coff_i, cnbr_i, k, l, 1);
        free(l);
    }
    double t4 = now_s();
    t[0] = t1 - t0;
    t[1] = t2 - t1;
    t[2] = t4 - t3;
    free(km);
    free(cnt);
    free(start);
    free(order);
    free(rank);
    free(vk);
    free(ek);
    free(coff_o);
    free(coff_i);
    free(cnbr_o);
    free(cnbr_i);
    return kmax;
}
