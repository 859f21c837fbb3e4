/*
 * oracle.c -- CPU restatement of the reference's parallel-A* hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Graph layout used by every function here is the reference's own file
 * layout (Appendix A of the reference's formats, GA_star/GenerateCSR.cpp:92-132):
 *   off[N]  row offsets, off[0] = 0, NO sentinel: the row of node N-1 ends at E
 *   col[E]  target ids, a negative id is a tombstone (GA_star/GA_v2.cu:122-125)
 *   w[E]    unsigned weights
 *
 * Functions and what pins them (tests/test_oracle.py):
 *   orc_astar_ref   GA_star/A_star_cpu.cpp:10-204. Pinned stdout-for-stdout
 *                   against the reference's own A_star_cpu.cpp compiled from
 *                   /root/reference into oracle/_ref (ties and the row-end
 *                   quirk included), and against orc_dijkstra.
 *   orc_dijkstra    textbook Dijkstra. Pinned by Floyd-Warshall brute force on
 *                   tiny graphs and by closed forms (open grid = Manhattan,
 *                   chains, s == t, unreachable).
 *   orc_bfs_hops    GA_star/bfs_hx.cu:11-36,158-172 (level-synchronous BFS,
 *                   -1 = unreachable). Pinned by brute-force hop counts
 *                   (Floyd-Warshall on unit weights) and closed forms.
 *   orc_transpose   DA_star/genOpCSR.cpp:31-165. Pinned byte-for-byte against
 *                   the reference's genOpCSR.cpp compiled into oracle/_ref.
 *   orc_csr_from_edges  GA_star/GenerateCSR.cpp:36-132. Pinned byte-for-byte
 *                   against the reference's GenerateCSR.cpp in oracle/_ref.
 *   orc_astar_batch_euclid  orc_astar_ref's search over many queries on a
 *                   thread pool, h = floor(Euclid) from coordinates. Pinned by
 *                   cost equality with orc_astar_ref on per-target h arrays.
 *   orc_apply_batch DA_star/GA_merge.cu:1019-1089 (removeDelEdges + mergeDiff):
 *                   G_{b+1} = G_b minus the first live instance of each deleted
 *                   (u,v) + inserts appended per source in arrival order.
 *                   Pinned against the reference's genCSRwholegraph.cpp
 *                   (DA_star/genCSRwholegraph.cpp:90-124) compiled into
 *                   oracle/_ref, on update files that list deletes first.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <pthread.h>

#define ORC_API __attribute__((visibility("default")))

static inline int row_end(const int *off, int N, int E, int u, int quirk)
{
    /* GA_star/A_star_cpu.cpp:169-172: the CPU oracle ends the row of node
     * N-1 at E-1 (dropping the last edge); the GPU kernels end it at E
     * (GA_star/GA_v2.cu:110-112). */
    if (u != N - 1) return off[u + 1];
    return quirk ? E - 1 : E;
}

/* ---- binary heap of node ids ordered by int cost, A_star_cpu.cpp:10-97 ---- */

static void ref_insert(int *pq, int *size, int node, const int *cx)
{
    /* A_star_cpu.cpp:10-31 */
    pq[*size] = node;
    (*size)++;
    int i = *size - 1;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (cx[pq[p]] > cx[pq[i]]) {
            int tmp = pq[i]; pq[i] = pq[p]; pq[p] = tmp;
            i = p;
        } else {
            break;
        }
    }
}

static int ref_getmin(int *pq, int *size, const int *cx)
{
    /* A_star_cpu.cpp:33-74, including its exact tie rules */
    if (*size == 0) return -1;
    int top = pq[0];
    pq[0] = pq[*size - 1];
    (*size)--;
    int i = 0;
    while (2 * i + 1 < *size) {
        int l = 2 * i + 1, r = 2 * i + 2;
        if (r >= *size) {
            if (cx[pq[l]] < cx[pq[i]]) {
                int tmp = pq[l]; pq[l] = pq[i]; pq[i] = tmp;
                i = l;
            } else {
                break;
            }
        } else if (cx[pq[i]] > cx[pq[l]] && cx[pq[l]] <= cx[pq[r]]) {
            int tmp = pq[l]; pq[l] = pq[i]; pq[i] = tmp;
            i = l;
        } else if (cx[pq[i]] > cx[pq[r]] && cx[pq[r]] <= cx[pq[l]]) {
            int tmp = pq[r]; pq[r] = pq[i]; pq[i] = tmp;
            i = r;
        } else {
            break;
        }
    }
    return top;
}

static void ref_update(int *pq, const int *size, int node, const int *cx)
{
    /* A_star_cpu.cpp:77-97: linear scan for the LAST occurrence, then sift up
     * only when the index is > 0. */
    int index = -1;
    for (int i = 0; i < *size; i++)
        if (pq[i] == node) index = i;
    if (index > 0) {
        int i = index;
        while (i > 0) {
            int p = (i - 1) / 2;
            if (cx[pq[p]] > cx[pq[i]]) {
                int tmp = pq[i]; pq[i] = pq[p]; pq[p] = tmp;
                i = p;
            } else {
                break;
            }
        }
    }
}

/* h(v): the given array, or floor(sqrt(dx^2 + dy^2)) to (tx, ty) computed
 * on demand from integer coordinates (exact: the float root is corrected). */
static inline int hval(const int *h, const int *xy, int tx, int ty, int v)
{
    if (h) return h[v];
    const int64_t dx = (int64_t)xy[2 * (size_t)v] - tx, dy = (int64_t)xy[2 * (size_t)v + 1] - ty;
    const int64_t d2 = dx * dx + dy * dy;
    int64_t r = (int64_t)__builtin_sqrt((double)d2);
    while (r * r > d2) r--;
    while ((r + 1) * (r + 1) <= d2) r++;
    return (int)r;
}

/* The search of A_star_cpu.cpp:100-204 on caller-provided buffers (cx, open,
 * pq of N+1 ints, parent), with the heuristic from an array or coordinates. */
static int astar_ref_core(int N, int E, const int *off, const int *col, const uint32_t *w,
                          const int *h, const int *xy, int s, int t, int quirk, int *cx, int *open,
                          int *pq, int *parent, int *cost_out, long long *expansions_out)
{
    int tx = 0, ty = 0;
    if (!h) { tx = xy[2 * (size_t)t]; ty = xy[2 * (size_t)t + 1]; }
    int size = 0;
    long long expanded = 0;
    for (int i = 0; i < N; i++) { cx[i] = INT_MAX; parent[i] = -1; open[i] = -1; }

    cx[s] = hval(h, xy, tx, ty, s);
    ref_insert(pq, &size, s, cx);
    open[s] = 1;
    int found = 0;
    while (size > 0 && !found) {
        int u = ref_getmin(pq, &size, cx);
        open[u] = -1;
        if (u == t) { found = 1; break; }
        expanded++;
        const int hu = hval(h, xy, tx, ty, u);
        int b = off[u], e = row_end(off, N, E, u, quirk);
        for (int j = b; j < e; j++) {
            int c = col[j];
            uint32_t cand = (uint32_t)(cx[u] - hu) + w[j] + (uint32_t)hval(h, xy, tx, ty, c);
            if ((uint32_t)cx[c] > cand) {
                cx[c] = (int)cand;
                if (open[c] == 1) {
                    ref_update(pq, &size, c, cx);
                } else {
                    /* pq holds at most one live copy per node, but reopened
                     * nodes may be pushed again: size never exceeds N. */
                    ref_insert(pq, &size, c, cx);
                    open[c] = 1;
                }
                parent[c] = u;
            }
        }
    }
    *cost_out = cx[t];
    if (expansions_out) *expansions_out = expanded;
    return found;
}

/*
 * A_star_cpu.cpp:100-204.  Returns 1 if the target was popped (found), else 0.
 * cost_out gets Cx[t] (f = g + h of the target); parent_out (N ints) gets the
 * parent array, so the reference's printed path t, parent[t], ... can be
 * rebuilt.  Relaxation arithmetic is unsigned 32-bit exactly as in
 * A_star_cpu.cpp:176 (int - int + unsigned + int); heap comparisons are signed
 * int as in A_star_cpu.cpp:21,45,55,61.
 */
ORC_API int orc_astar_ref(int N, int E, const int *off, const int *col,
                          const uint32_t *w, const int *h, int s, int t,
                          int quirk, int *cost_out, int *parent_out,
                          long long *expansions_out)
{
    int *cx = (int *)malloc(sizeof(int) * (size_t)N);
    int *open = (int *)malloc(sizeof(int) * (size_t)N);
    int *pq = (int *)malloc(sizeof(int) * ((size_t)N + 1));
    if (!cx || !open || !pq) { free(cx); free(open); free(pq); return -1; }
    int found = astar_ref_core(N, E, off, col, w, h, NULL, s, t, quirk, cx, open, pq, parent_out, cost_out,
                               expansions_out);
    free(cx); free(open); free(pq);
    return found;
}

/* ---- many queries of A_star_cpu.cpp's search on a thread pool ----
 * The CPU baseline of the batched configurations: nthreads workers take
 * queries from a shared counter, one query per worker at a time, each with
 * its own N-sized buffers (re-initialised per query, as the reference's one
 * process per query does).  h is floor(Euclid) to each query's target,
 * computed on demand from the coordinates.  Pinned by equality of every cost
 * with orc_astar_ref on a per-target h array (tests/test_oracle.py). */
typedef struct {
    int N, E;
    const int *off, *col, *xy;
    const uint32_t *w;
    int nq;
    const int *src, *dst;
    int *costs;
    long long expanded;
    int next;
    int failed;
    pthread_mutex_t mu;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *J = (batch_job *)arg;
    int *cx = (int *)malloc(sizeof(int) * (size_t)J->N);
    int *open = (int *)malloc(sizeof(int) * (size_t)J->N);
    int *pq = (int *)malloc(sizeof(int) * ((size_t)J->N + 1));
    int *parent = (int *)malloc(sizeof(int) * (size_t)J->N);
    long long mine = 0;
    if (!cx || !open || !pq || !parent) {
        pthread_mutex_lock(&J->mu);
        J->failed = 1;
        pthread_mutex_unlock(&J->mu);
    } else {
        for (;;) {
            int q = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
            if (q >= J->nq) break;
            long long ex = 0;
            int cost = 0;
            int found = astar_ref_core(J->N, J->E, J->off, J->col, J->w, NULL, J->xy, J->src[q], J->dst[q], 0, cx,
                                       open, pq, parent, &cost, &ex);
            J->costs[q] = found ? cost : INT_MAX;
            mine += ex;
        }
    }
    free(cx); free(open); free(pq); free(parent);
    pthread_mutex_lock(&J->mu);
    J->expanded += mine;
    pthread_mutex_unlock(&J->mu);
    return NULL;
}

ORC_API int orc_astar_batch_euclid(int N, int E, const int *off, const int *col, const uint32_t *w,
                                   const int *xy, int nq, const int *src, const int *dst, int nthreads,
                                   int *costs_out, long long *expansions_out)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    batch_job J = {N, E, off, col, xy, w, nq, src, dst, costs_out, 0, 0, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    int started = 0;
    for (int i = 0; i < nthreads; i++)
        if (pthread_create(&th[started], NULL, batch_worker, &J) == 0) started++;
    if (started == 0) batch_worker(&J);
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
    if (expansions_out) *expansions_out = J.expanded;
    return J.failed ? -1 : started ? started : 1;
}

/* ---- textbook Dijkstra with an indexed binary heap (64-bit distances) ---- */

typedef struct { long long *key; int *heap; int *pos; int n; } ixheap;

static void ix_swap(ixheap *H, int a, int b)
{
    int na = H->heap[a], nb = H->heap[b];
    H->heap[a] = nb; H->heap[b] = na;
    H->pos[nb] = a; H->pos[na] = b;
}

static void ix_up(ixheap *H, int i)
{
    while (i > 0) {
        int p = (i - 1) / 2;
        if (H->key[H->heap[p]] > H->key[H->heap[i]]) { ix_swap(H, p, i); i = p; }
        else break;
    }
}

static void ix_down(ixheap *H, int i)
{
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        if (l < H->n && H->key[H->heap[l]] < H->key[H->heap[m]]) m = l;
        if (r < H->n && H->key[H->heap[r]] < H->key[H->heap[m]]) m = r;
        if (m == i) return;
        ix_swap(H, i, m);
        i = m;
    }
}

/*
 * Single-source shortest distances from s (all nodes, no early exit), with
 * tombstones (col < 0) skipped and the row of N-1 ending at E.  dist_out is
 * int64; unreachable = -1.  If stop_at >= 0 the search stops once stop_at is
 * settled (its distance is then final).  Returns the number of settled nodes.
 */
ORC_API long long orc_dijkstra(int N, int E, const int *off, const int *col,
                               const uint32_t *w, int s, int stop_at,
                               long long *dist_out)
{
    ixheap H;
    H.key = dist_out;
    H.heap = (int *)malloc(sizeof(int) * (size_t)N);
    H.pos = (int *)malloc(sizeof(int) * (size_t)N);
    H.n = 0;
    char *done = (char *)calloc((size_t)N, 1);
    long long settled = 0;
    for (int i = 0; i < N; i++) { dist_out[i] = LLONG_MAX; H.pos[i] = -1; }
    dist_out[s] = 0;
    H.heap[0] = s; H.pos[s] = 0; H.n = 1;
    while (H.n > 0) {
        int u = H.heap[0];
        ix_swap(&H, 0, H.n - 1);
        H.n--;
        H.pos[u] = -1;
        if (H.n > 0) ix_down(&H, 0);
        done[u] = 1;
        settled++;
        if (u == stop_at) break;
        int b = off[u], e = (u == N - 1) ? E : off[u + 1];
        for (int j = b; j < e; j++) {
            int c = col[j];
            if (c < 0 || done[c]) continue;
            long long nd = dist_out[u] + (long long)w[j];
            if (nd < dist_out[c]) {
                dist_out[c] = nd;
                if (H.pos[c] < 0) { H.heap[H.n] = c; H.pos[c] = H.n; H.n++; }
                ix_up(&H, H.pos[c]);
            }
        }
    }
    for (int i = 0; i < N; i++)
        if (dist_out[i] == LLONG_MAX) dist_out[i] = -1;
    free(H.heap); free(H.pos); free(done);
    return settled;
}

/* ---- the same A* with an indexed heap (labelled variant, CPU baseline) ----
 * A_star_cpu.cpp's search with its O(heap) linear-scan decrease-key replaced
 * by a position index (O(log heap)); same relax arithmetic, same stop rule
 * (first pop of t).  Costs equal orc_astar_ref's; heap positions of equal
 * keys may differ, so parent chains under ties may too.  Pinned by equality
 * of costs with orc_astar_ref and orc_dijkstra (tests/test_oracle.py). */
ORC_API int orc_astar_indexed(int N, int E, const int *off, const int *col, const uint32_t *w,
                              const int *h, int s, int t, int *cost_out, long long *expansions_out)
{
    long long *key = (long long *)malloc(sizeof(long long) * (size_t)N);
    int *cx = (int *)malloc(sizeof(int) * (size_t)N);
    int *heap = (int *)malloc(sizeof(int) * (size_t)N);
    int *pos = (int *)malloc(sizeof(int) * (size_t)N);
    if (!key || !cx || !heap || !pos) { free(key); free(cx); free(heap); free(pos); return -1; }
    for (int i = 0; i < N; i++) { cx[i] = INT_MAX; pos[i] = -1; }
    ixheap H = {key, heap, pos, 0};
    long long expanded = 0;
    int found = 0;
    cx[s] = h[s];
    key[s] = cx[s];
    heap[0] = s; pos[s] = 0; H.n = 1;
    while (H.n > 0) {
        int u = heap[0];
        ix_swap(&H, 0, H.n - 1);
        H.n--;
        pos[u] = -1;
        if (H.n > 0) ix_down(&H, 0);
        if (u == t) { found = 1; break; }
        expanded++;
        int b = off[u], e = (u == N - 1) ? E : off[u + 1];
        for (int j = b; j < e; j++) {
            int c = col[j];
            if (c < 0) continue;
            uint32_t cand = (uint32_t)(cx[u] - h[u]) + w[j] + (uint32_t)h[c];
            if ((uint32_t)cx[c] > cand) {
                cx[c] = (int)cand;
                key[c] = cx[c];
                if (pos[c] < 0) { heap[H.n] = c; pos[c] = H.n; H.n++; }
                ix_up(&H, pos[c]);
            }
        }
    }
    *cost_out = cx[t];
    if (expansions_out) *expansions_out = expanded;
    free(key); free(cx); free(heap); free(pos);
    return found;
}

/*
 * Level-synchronous BFS from t over the graph given (the reference feeds the
 * REVERSED graph: GA_star/bfs_hx.cu:45,111-138), tombstones skipped.
 * dist_out[v] = hop count, -1 if unreachable (bfs_hx.cu:72,163,170).
 * Returns the number of levels (max hop count + 1), 0 if t is out of range.
 */
ORC_API int orc_bfs_hops(int N, int E, const int *off, const int *col, int t,
                         int *dist_out)
{
    for (int i = 0; i < N; i++) dist_out[i] = -1;
    if (t < 0 || t >= N) return 0;
    int *cur = (int *)malloc(sizeof(int) * (size_t)N);
    int *nxt = (int *)malloc(sizeof(int) * (size_t)N);
    int nc = 1, levels = 0;
    cur[0] = t;
    dist_out[t] = 0;
    while (nc > 0) {
        int nn = 0;
        levels++;
        for (int k = 0; k < nc; k++) {
            int u = cur[k];
            int b = off[u], e = (u == N - 1) ? E : off[u + 1];
            for (int j = b; j < e; j++) {
                int c = col[j];
                if (c < 0) continue;
                if (dist_out[c] < 0) { dist_out[c] = dist_out[u] + 1; nxt[nn++] = c; }
            }
        }
        int *tmp = cur; cur = nxt; nxt = tmp;
        nc = nn;
    }
    free(cur); free(nxt);
    return levels;
}

/*
 * Transpose in DA_star/genOpCSR.cpp:55-74 order: sources are visited in id
 * order and each edge u->v (weight w) is appended to v's reverse row as (u, w),
 * so reverse rows list predecessors by ascending source id, and parallel
 * edges keep their forward order.  Tombstones are dropped only if drop_dead.
 * roff has N entries (reference layout); returns the reverse edge count.
 */
ORC_API int orc_transpose(int N, int E, const int *off, const int *col,
                          const uint32_t *w, int drop_dead, int *roff,
                          int *rcol, uint32_t *rw)
{
    int *cnt = (int *)calloc((size_t)N + 1, sizeof(int));
    for (int u = 0; u < N; u++) {
        int b = off[u], e = (u == N - 1) ? E : off[u + 1];
        for (int j = b; j < e; j++) if (col[j] >= 0) cnt[col[j] + 1]++;
    }
    for (int i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    int RE = cnt[N];
    for (int i = 0; i < N; i++) roff[i] = cnt[i];
    for (int u = 0; u < N; u++) {
        int b = off[u], e = (u == N - 1) ? E : off[u + 1];
        for (int j = b; j < e; j++) {
            int c = col[j];
            if (c < 0) continue;
            int p = cnt[c]++;
            rcol[p] = u;
            rw[p] = w[j];
        }
    }
    (void)drop_dead;
    free(cnt);
    return RE;
}

/*
 * Edge triples -> CSR as GA_star/GenerateCSR.cpp:36-132 writes it:
 * N = max id + 1, rows in id order, each row in arrival order.
 * off gets N entries.  Returns N.
 */
ORC_API int orc_csr_from_edges(int M, const int *us, const int *vs,
                               const uint32_t *ws, int *off, int *col,
                               uint32_t *w)
{
    int N = 0;
    for (int i = 0; i < M; i++) {
        if (us[i] + 1 > N) N = us[i] + 1;
        if (vs[i] + 1 > N) N = vs[i] + 1;
    }
    int *cnt = (int *)calloc((size_t)N + 1, sizeof(int));
    for (int i = 0; i < M; i++) cnt[us[i] + 1]++;
    for (int i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    for (int i = 0; i < N; i++) off[i] = cnt[i];
    for (int i = 0; i < M; i++) {
        int p = cnt[us[i]]++;
        col[p] = vs[i];
        w[p] = ws[i];
    }
    free(cnt);
    return N;
}

/*
 * One update batch, DA_star/GA_merge.cu:696-933 semantics:
 *  1. for each delete (flag 0) in file order: tombstone the first live
 *     instance of v in u's row (removeDelEdges, GA_merge.cu:1019-1037); a
 *     delete with no live match is a no-op;
 *  2. merge (mergeDiff, GA_merge.cu:1042-1089): per node, surviving edges in
 *     row order, then that node's inserts (flag 1) in file order.
 * Output arrays must hold E + (#inserts) entries (ooff: N).  Returns the new
 * edge count; *n_deleted gets the number of effective deletes.
 * The input arrays are not modified.
 */
ORC_API int orc_apply_batch(int N, int E, const int *off, const int *col,
                            const uint32_t *w, int nup, const int *flag,
                            const int *us, const int *vs, const uint32_t *ws,
                            int *ooff, int *ocol, uint32_t *ow, int *n_deleted)
{
    int *c2 = (int *)malloc(sizeof(int) * ((size_t)E + 1));
    memcpy(c2, col, sizeof(int) * (size_t)E);
    int ndel = 0, nins = 0;
    for (int i = 0; i < nup; i++) {
        if (flag[i] == 0) {
            int u = us[i], v = vs[i];
            int b = off[u], e = (u == N - 1) ? E : off[u + 1];
            for (int j = b; j < e; j++)
                if (c2[j] == v) { c2[j] = -1; ndel++; break; }
        } else if (flag[i] == 1) {
            nins++;
        }
    }
    /* inserts grouped by source, stable */
    int *icnt = (int *)calloc((size_t)N + 1, sizeof(int));
    int *iord = (int *)malloc(sizeof(int) * ((size_t)nins + 1));
    for (int i = 0; i < nup; i++) if (flag[i] == 1) icnt[us[i] + 1]++;
    for (int i = 0; i < N; i++) icnt[i + 1] += icnt[i];
    {
        int *fill = (int *)malloc(sizeof(int) * ((size_t)N + 1));
        memcpy(fill, icnt, sizeof(int) * ((size_t)N + 1));
        for (int i = 0; i < nup; i++) if (flag[i] == 1) iord[fill[us[i]]++] = i;
        free(fill);
    }
    int k = 0;
    for (int u = 0; u < N; u++) {
        ooff[u] = k;
        int b = off[u], e = (u == N - 1) ? E : off[u + 1];
        for (int j = b; j < e; j++)
            if (c2[j] >= 0) { ocol[k] = c2[j]; ow[k] = w[j]; k++; }
        for (int q = icnt[u]; q < icnt[u + 1]; q++) {
            int i = iord[q];
            ocol[k] = vs[i]; ow[k] = ws[i]; k++;
        }
    }
    free(c2); free(icnt); free(iord);
    if (n_deleted) *n_deleted = ndel;
    return k;
}
