// Host-side text formats, host CSR handles and the update-batch semantics.
#include <algorithm>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/ddp_capi.h"
#include "../../include/ddp_io.h"
#include "hgraph.h"

static void put_err(char *err, int len, const char *fmt, ...)
{
    if (!err || len <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)len, fmt, ap);
    va_end(ap);
}

extern "C" int ddp_hgraph_from_arrays(int32_t N, int32_t E, const int32_t *off, const int32_t *col,
                                      const uint32_t *w, ddp_hgraph **out)
{
    if (!out || N <= 0 || E < 0 || !off || (E && (!col || !w))) return DDP_ERR_ARG;
    ddp_hgraph *g = new ddp_hgraph();
    g->N = N;
    g->E = E;
    g->off.assign(off, off + N);
    g->col.assign(col, col + E);
    g->w.assign(w, w + E);
    *out = g;
    return DDP_OK;
}

extern "C" void ddp_hgraph_free(ddp_hgraph *g) { delete g; }

extern "C" int ddp_hgraph_view(const ddp_hgraph *g, int32_t *N, int32_t *E, const int32_t **off,
                               const int32_t **col, const uint32_t **w)
{
    if (!g) return DDP_ERR_ARG;
    if (N) *N = g->N;
    if (E) *E = g->E;
    if (off) *off = g->off.data();
    if (col) *col = g->col.data();
    if (w) *w = g->w.data();
    return DDP_OK;
}

extern "C" const int32_t *ddp_hgraph_coords(const ddp_hgraph *g)
{
    return (g && !g->xy.empty()) ? g->xy.data() : nullptr;
}

extern "C" int ddp_hgraph_set_coords(ddp_hgraph *g, const int32_t *xy)
{
    if (!g || !xy) return DDP_ERR_ARG;
    g->xy.assign(xy, xy + 2 * (size_t)g->N);
    return DDP_OK;
}

// ------------------------------------------------------------ tokenizer

namespace {

struct MappedFile {
    const char *p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~MappedFile()
    {
        if (p && n) munmap((void *)p, n);
        if (fd >= 0) close(fd);
    }
    bool open_(const char *path)
    {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) != 0) return false;
        n = (size_t)st.st_size;
        if (n == 0) return true;
        void *m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            p = nullptr;
            n = 0;
            return false;
        }
        p = (const char *)m;
        return true;
    }
};

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Parses whitespace-separated integers of [b, e) into `sink(index, value)`,
// starting at token index `first`.  Returns the number of tokens, or -1 on a
// malformed token.
template <class Sink>
int64_t parse_range(const char *b, const char *e, int64_t first, Sink &&sink)
{
    int64_t k = 0;
    const char *s = b;
    while (s < e) {
        while (s < e && is_space(*s)) ++s;
        if (s >= e) break;
        bool neg = false;
        if (*s == '-' || *s == '+') {
            neg = *s == '-';
            ++s;
        }
        if (s >= e || *s < '0' || *s > '9') return -1;
        int64_t v = 0;
        while (s < e && *s >= '0' && *s <= '9') {
            v = v * 10 + (*s - '0');
            if (v > (int64_t)1 << 40) return -1;
            ++s;
        }
        if (s < e && !is_space(*s)) return -1;
        sink(first + k, neg ? -v : v);
        ++k;
    }
    return k;
}

// Parallel parse of all integer tokens of [b, e): chunk boundaries are moved
// to whitespace, each thread counts its tokens, then parses into place.
template <class Sink>
int64_t parse_parallel(const char *b, const char *e, Sink &&sink)
{
    const size_t n = (size_t)(e - b);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < (1u << 20)) nt = 1;
    std::vector<const char *> cut(nt + 1);
    cut[0] = b;
    cut[nt] = e;
    for (unsigned i = 1; i < nt; ++i) {
        const char *c = b + n * i / nt;
        if (c < cut[i - 1]) c = cut[i - 1];
        while (c < e && !is_space(*c)) ++c;
        cut[i] = c;
    }
    std::vector<int64_t> cnt(nt, 0);
    {
        std::vector<std::thread> th;
        for (unsigned i = 0; i < nt; ++i)
            th.emplace_back([&, i] { cnt[i] = parse_range(cut[i], cut[i + 1], 0, [](int64_t, int64_t) {}); });
        for (auto &t : th) t.join();
    }
    int64_t total = 0;
    std::vector<int64_t> start(nt);
    for (unsigned i = 0; i < nt; ++i) {
        if (cnt[i] < 0) return -1;
        start[i] = total;
        total += cnt[i];
    }
    {
        std::vector<std::thread> th;
        for (unsigned i = 0; i < nt; ++i)
            th.emplace_back([&, i] { parse_range(cut[i], cut[i + 1], start[i], sink); });
        for (auto &t : th) t.join();
    }
    return total;
}

}  // namespace

extern "C" int ddp_read_graph_txt(const char *path, ddp_hgraph **out, char *err, int errlen)
{
    if (!path || !out) return DDP_ERR_ARG;
    *out = nullptr;
    MappedFile f;
    if (!f.open_(path)) {
        put_err(err, errlen, "cannot open %s: %s", path, strerror(errno));
        return DDP_ERR_IO;
    }
    // header "N E"
    int64_t hdr[2] = {0, 0};
    const char *s = f.p, *e = f.p + f.n;
    int got = 0;
    const char *body = s;
    while (got < 2 && body < e) {
        while (body < e && is_space(*body)) ++body;
        const char *t = body;
        while (t < e && !is_space(*t)) ++t;
        if (parse_range(body, t, got, [&](int64_t i, int64_t v) { hdr[i] = v; }) != 1) break;
        ++got;
        body = t;
    }
    if (got != 2 || hdr[0] <= 0 || hdr[1] < 0 || hdr[0] > INT32_MAX || hdr[1] > INT32_MAX) {
        put_err(err, errlen, "%s: bad header (expected \"N E\")", path);
        return DDP_ERR_FORMAT;
    }
    const int32_t N = (int32_t)hdr[0], E = (int32_t)hdr[1];
    ddp_hgraph *g = new ddp_hgraph();
    g->N = N;
    g->E = E;
    g->col.resize(E);
    g->off.resize(N);
    g->w.resize(E);
    const int64_t want = 2 * (int64_t)E + N;
    const int64_t got_tokens = parse_parallel(body, e, [&](int64_t i, int64_t v) {
        if (i < E) g->col[i] = (int32_t)v;
        else if (i < (int64_t)E + N) g->off[i - E] = (int32_t)v;
        else if (i < want) g->w[i - E - N] = (uint32_t)v;
    });
    if (got_tokens < 0) {
        delete g;
        put_err(err, errlen, "%s: malformed integer token", path);
        return DDP_ERR_FORMAT;
    }
    if (got_tokens < want) {
        delete g;
        put_err(err, errlen, "%s: truncated (%lld of %lld numbers after the header)", path, (long long)got_tokens,
                (long long)want);
        return DDP_ERR_FORMAT;
    }
    for (int32_t i = 0; i < N; ++i) {
        const int32_t end = (i + 1 < N) ? g->off[i + 1] : E;
        if (g->off[i] < 0 || g->off[i] > end || end > E) {
            put_err(err, errlen, "%s: row %d offsets [%d,%d) invalid", path, i, g->off[i], end);
            delete g;
            return DDP_ERR_FORMAT;
        }
    }
    for (int32_t j = 0; j < E; ++j)
        if (g->col[j] >= N) {
            put_err(err, errlen, "%s: column %d = %d >= N", path, j, g->col[j]);
            delete g;
            return DDP_ERR_FORMAT;
        }
    *out = g;
    return DDP_OK;
}

namespace {

struct Writer {
    FILE *fp;
    std::vector<char> buf;
    size_t n = 0;
    explicit Writer(FILE *f) : fp(f), buf(1 << 22) {}
    void flush()
    {
        if (n) fwrite(buf.data(), 1, n, fp);
        n = 0;
    }
    void room(size_t k)
    {
        if (n + k > buf.size()) flush();
    }
    void num(int64_t v)
    {
        room(24);
        char tmp[24];
        int k = 0;
        const bool neg = v < 0;
        uint64_t u = neg ? (uint64_t)(-v) : (uint64_t)v;
        do {
            tmp[k++] = (char)('0' + u % 10);
            u /= 10;
        } while (u);
        if (neg) buf[n++] = '-';
        while (k) buf[n++] = tmp[--k];
    }
    void ch(char c)
    {
        room(1);
        buf[n++] = c;
    }
};

}  // namespace

extern "C" int ddp_write_graph_txt(const char *path, const ddp_hgraph *g)
{
    if (!path || !g) return DDP_ERR_ARG;
    FILE *fp = fopen(path, "w");
    if (!fp) return DDP_ERR_IO;
    Writer w(fp);
    // token layout of GA_star/GenerateCSR.cpp:92-132: "N E\n", each value
    // followed by a space, a newline after each section.
    w.num(g->N);
    w.ch(' ');
    w.num(g->E);
    w.ch('\n');
    for (int32_t j = 0; j < g->E; ++j) {
        w.num(g->col[j]);
        w.ch(' ');
    }
    w.ch('\n');
    for (int32_t i = 0; i < g->N; ++i) {
        w.num(g->off[i]);
        w.ch(' ');
    }
    w.ch('\n');
    for (int32_t j = 0; j < g->E; ++j) {
        w.num(g->w[j]);
        w.ch(' ');
    }
    w.ch('\n');
    w.flush();
    const bool ok = ferror(fp) == 0;
    fclose(fp);
    return ok ? DDP_OK : DDP_ERR_IO;
}

extern "C" int ddp_read_hx_txt(const char *path, int32_t n, int32_t *h, char *err, int errlen)
{
    if (!path || !h || n <= 0) return DDP_ERR_ARG;
    MappedFile f;
    if (!f.open_(path)) {
        put_err(err, errlen, "cannot open %s: %s", path, strerror(errno));
        return DDP_ERR_IO;
    }
    const int64_t got = parse_parallel(f.p, f.p + f.n, [&](int64_t i, int64_t v) {
        if (i < n) h[i] = (int32_t)v;
    });
    if (got < 0) {
        put_err(err, errlen, "%s: malformed integer token", path);
        return DDP_ERR_FORMAT;
    }
    if (got < n) {
        put_err(err, errlen, "%s: truncated (%lld of %d values)", path, (long long)got, n);
        return DDP_ERR_FORMAT;
    }
    return DDP_OK;
}

extern "C" int ddp_write_hx_txt(const char *path, int32_t n, const int32_t *h)
{
    if (!path || !h || n < 0) return DDP_ERR_ARG;
    FILE *fp = fopen(path, "w");
    if (!fp) return DDP_ERR_IO;
    Writer w(fp);
    for (int32_t i = 0; i < n; ++i) {  // "%d " per value (GA_star/bfs_hx.cu:170)
        w.num(h[i]);
        w.ch(' ');
    }
    w.flush();
    const bool ok = ferror(fp) == 0;
    fclose(fp);
    return ok ? DDP_OK : DDP_ERR_IO;
}

extern "C" int ddp_hgraph_transpose(const ddp_hgraph *g, ddp_hgraph **out)
{
    if (!g || !out) return DDP_ERR_ARG;
    const int32_t N = g->N, E = g->E;
    ddp_hgraph *r = new ddp_hgraph();
    r->N = N;
    std::vector<int32_t> cnt(N + 1, 0);
    for (int32_t j = 0; j < E; ++j)
        if (g->col[j] >= 0) cnt[g->col[j] + 1]++;
    for (int32_t i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
    r->E = cnt[N];
    r->off.assign(cnt.begin(), cnt.end() - 1);
    r->col.resize(r->E);
    r->w.resize(r->E);
    for (int32_t u = 0; u < N; ++u) {
        const int32_t b = g->off[u], e = (u + 1 < N) ? g->off[u + 1] : E;
        for (int32_t j = b; j < e; ++j) {
            const int32_t c = g->col[j];
            if (c < 0) continue;
            const int32_t p = cnt[c]++;
            r->col[p] = u;
            r->w[p] = g->w[j];
        }
    }
    *out = r;
    return DDP_OK;
}

// ---------------------------------------------------------------- updates

extern "C" int ddp_read_updates_txt(const char *path, ddp_updates **out, char *err, int errlen)
{
    if (!path || !out) return DDP_ERR_ARG;
    *out = nullptr;
    MappedFile f;
    if (!f.open_(path)) {
        put_err(err, errlen, "cannot open %s: %s", path, strerror(errno));
        return DDP_ERR_IO;
    }
    std::vector<int64_t> tok;
    const int64_t got = parse_range(f.p, f.p + f.n, 0, [&](int64_t, int64_t v) { tok.push_back(v); });
    if (got < 0) {
        put_err(err, errlen, "%s: malformed integer token", path);
        return DDP_ERR_FORMAT;
    }
    ddp_updates *u = new ddp_updates();
    u->start.clear();
    size_t i = 0;
    while (i < tok.size()) {
        const int64_t c = tok[i++];
        if (c < 0 || i + 4 * (size_t)c > tok.size()) {
            put_err(err, errlen, "%s: batch %zu truncated or bad count %lld", path, u->start.size(), (long long)c);
            delete u;
            return DDP_ERR_FORMAT;
        }
        u->start.push_back((int32_t)u->lines.size());
        for (int64_t k = 0; k < c; ++k, i += 4) {
            ddp_update_t l{(int32_t)tok[i], (int32_t)tok[i + 1], (int32_t)tok[i + 2], (uint32_t)tok[i + 3]};
            if (l.flag != 0 && l.flag != 1) {
                put_err(err, errlen, "%s: flag %d is neither 0 (delete) nor 1 (insert)", path, l.flag);
                delete u;
                return DDP_ERR_FORMAT;
            }
            u->lines.push_back(l);
        }
    }
    u->start.push_back((int32_t)u->lines.size());
    *out = u;
    return DDP_OK;
}

extern "C" int ddp_write_updates_txt(const char *path, const ddp_updates *u)
{
    if (!path || !u) return DDP_ERR_ARG;
    FILE *fp = fopen(path, "w");
    if (!fp) return DDP_ERR_IO;
    for (size_t b = 0; b + 1 < u->start.size(); ++b) {
        fprintf(fp, "%d\n", u->start[b + 1] - u->start[b]);
        for (int32_t k = u->start[b]; k < u->start[b + 1]; ++k) {
            const ddp_update_t &l = u->lines[k];
            fprintf(fp, "%d %d %d %u\n", l.flag, l.u, l.v, l.w);
        }
    }
    const bool ok = ferror(fp) == 0;
    fclose(fp);
    return ok ? DDP_OK : DDP_ERR_IO;
}

extern "C" int ddp_updates_count(const ddp_updates *u) { return u ? (int)u->start.size() - 1 : 0; }

extern "C" int ddp_updates_batch(const ddp_updates *u, int32_t b, const ddp_update_t **lines, int32_t *n)
{
    if (!u || b < 0 || b + 1 >= (int32_t)u->start.size() || !lines || !n) return DDP_ERR_ARG;
    *lines = u->lines.data() + u->start[b];
    *n = u->start[b + 1] - u->start[b];
    return DDP_OK;
}

extern "C" void ddp_updates_free(ddp_updates *u) { delete u; }

extern "C" int ddp_hgraph_apply_batch(const ddp_hgraph *g, const ddp_update_t *lines, int32_t n,
                                      ddp_hgraph **out, int32_t *n_deleted)
{
    if (!g || !out || (n && !lines)) return DDP_ERR_ARG;
    const int32_t N = g->N, E = g->E;
    std::vector<int32_t> c2(g->col);
    int32_t ndel = 0;
    std::vector<int32_t> icnt(N + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        const ddp_update_t &l = lines[i];
        if (l.u < 0 || l.u >= N || l.v < 0 || l.v >= N) return DDP_ERR_ARG;
        if (l.flag == 0) {
            const int32_t b = g->off[l.u], e = (l.u + 1 < N) ? g->off[l.u + 1] : E;
            for (int32_t j = b; j < e; ++j)
                if (c2[j] == l.v) {
                    c2[j] = -1;
                    ++ndel;
                    break;
                }
        } else {
            icnt[l.u + 1]++;
        }
    }
    for (int32_t i = 0; i < N; ++i) icnt[i + 1] += icnt[i];
    std::vector<int32_t> ord(icnt[N]);
    {
        std::vector<int32_t> fill(icnt.begin(), icnt.end() - 1);
        for (int32_t i = 0; i < n; ++i)
            if (lines[i].flag == 1) ord[fill[lines[i].u]++] = i;
    }
    ddp_hgraph *r = new ddp_hgraph();
    r->N = N;
    r->off.resize(N);
    r->col.reserve((size_t)E + icnt[N]);
    r->w.reserve((size_t)E + icnt[N]);
    for (int32_t u = 0; u < N; ++u) {
        r->off[u] = (int32_t)r->col.size();
        const int32_t b = g->off[u], e = (u + 1 < N) ? g->off[u + 1] : E;
        for (int32_t j = b; j < e; ++j)
            if (c2[j] >= 0) {
                r->col.push_back(c2[j]);
                r->w.push_back(g->w[j]);
            }
        for (int32_t k = icnt[u]; k < icnt[u + 1]; ++k) {
            r->col.push_back(lines[ord[k]].v);
            r->w.push_back(lines[ord[k]].w);
        }
    }
    r->E = (int32_t)r->col.size();
    r->xy = g->xy;
    if (n_deleted) *n_deleted = ndel;
    *out = r;
    return DDP_OK;
}

// ------------------------------------------------------------ binary cache
//
// An exact binary image of a parsed graph.txt: a header naming the text file's
// size and modification time, then off[N], col[E], w[E] as stored in memory.
// A cache whose header does not match the text file (or is unreadable) is
// ignored and rewritten; failing to write one is not an error.

namespace {

struct CacheHeader {
    char magic[8];            // "DDPGRPH1"
    uint64_t src_size;
    int64_t src_mtime_s, src_mtime_ns;
    int32_t N, E;
};

bool text_identity(const char *path, uint64_t *size, int64_t *sec, int64_t *nsec)
{
    struct stat st;
    if (stat(path, &st) != 0) return false;
    *size = (uint64_t)st.st_size;
    *sec = (int64_t)st.st_mtim.tv_sec;
    *nsec = (int64_t)st.st_mtim.tv_nsec;
    return true;
}

bool read_cache(const char *cache, const CacheHeader &want, ddp_hgraph **out)
{
    MappedFile f;
    if (!f.open_(cache) || f.n < sizeof(CacheHeader)) return false;
    CacheHeader h;
    std::memcpy(&h, f.p, sizeof h);
    if (std::memcmp(h.magic, "DDPGRPH1", 8) || h.src_size != want.src_size || h.src_mtime_s != want.src_mtime_s ||
        h.src_mtime_ns != want.src_mtime_ns || h.N <= 0 || h.E < 0)
        return false;
    const size_t need = sizeof h + (size_t)h.N * 4 + (size_t)h.E * 8;
    if (f.n != need) return false;
    ddp_hgraph *g = new ddp_hgraph();
    g->N = h.N;
    g->E = h.E;
    const char *p = f.p + sizeof h;
    g->off.resize(h.N);
    g->col.resize(h.E);
    g->w.resize(h.E);
    std::memcpy(g->off.data(), p, (size_t)h.N * 4);
    std::memcpy(g->col.data(), p + (size_t)h.N * 4, (size_t)h.E * 4);
    std::memcpy(g->w.data(), p + (size_t)h.N * 4 + (size_t)h.E * 4, (size_t)h.E * 4);
    *out = g;
    return true;
}

void write_cache(const char *cache, CacheHeader h, const ddp_hgraph *g)
{
    const std::string tmp = std::string(cache) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    std::memcpy(h.magic, "DDPGRPH1", 8);
    h.N = g->N;
    h.E = g->E;
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(g->off.data(), 4, g->off.size(), f) == g->off.size() &&
              std::fwrite(g->col.data(), 4, g->col.size(), f) == g->col.size() &&
              std::fwrite(g->w.data(), 4, g->w.size(), f) == g->w.size();
    ok = (std::fclose(f) == 0) && ok;
    if (ok) ok = std::rename(tmp.c_str(), cache) == 0;  // readers never see a partial cache
    if (!ok) std::remove(tmp.c_str());
}

}  // namespace

extern "C" int ddp_read_graph_cached(const char *path, const char *cache_path, ddp_hgraph **out, char *err,
                                     int errlen)
{
    if (!path || !out) return DDP_ERR_ARG;
    *out = nullptr;
    const std::string cache = cache_path ? std::string(cache_path) : std::string(path) + ".ddpbin";
    CacheHeader want{};
    if (!text_identity(path, &want.src_size, &want.src_mtime_s, &want.src_mtime_ns)) {
        put_err(err, errlen, "cannot open %s: %s", path, strerror(errno));
        return DDP_ERR_IO;
    }
    if (read_cache(cache.c_str(), want, out)) return DDP_OK;
    const int rc = ddp_read_graph_txt(path, out, err, errlen);
    if (rc == DDP_OK) write_cache(cache.c_str(), want, *out);
    return rc;
}
