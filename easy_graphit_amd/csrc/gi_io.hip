// Graph files, read on the host: the reference's load_graph
// (runtime/infra_gpu/graph.h:236-303) together with the GAPBS reader and
// builder it calls (runtime/infra_gapbs/reader.h:51-375, builder.h:62-246).
//
// Formats by suffix (reader.h:303-330):
//   .el          "u v" pairs; weight 1                      (ReadInEL, reader.h:51-58)
//   .wel         "u v w" triples                            (ReadInWEL, :60-68)
//   .gr          DIMACS: lines "a u v w"; ids as written    (ReadInGR, :70-85)
//   .graph       METIS: header "n m [fmt]", one adjacency
//                line per vertex, ids from 1; fmt 1 = weights (ReadInMetis, :88-141)
//   .mtx         MatrixMarket coordinate, ids from 1; a symmetric matrix
//                adds every entry both ways; integer/real values are the
//                weights as an int32 read of the text                (ReadInMTX, :145-213)
//   .sg / .wsg   GAPBS serialized CSR, taken as stored      (ReadSerializedGraph, :332-373)
//   .graphit_bin the reference's cache of a loaded graph    (graph.h:252-268, 286-293)
// A text format becomes an edge list and then a CSR the way GAPBS's builder
// makes one: V = largest id + 1, each list sorted by (neighbour, weight), a
// repeated neighbour kept once with its least weight, self-loops dropped
// (builder.h:62-71, 134-163, 186-226). Without symmetrize the graph is
// directed, as in load_graph, which never symmetrizes (graph.h:273).
//
// Text is scanned the way the reference's stream reads (`in >> x` on int32):
// whitespace skipped, an optional sign, digits; the first token that does
// not read as an int32 ends an .el / .wel / .gr file, as the reference's read
// loops end. Where the reference reads an uninitialised value (a blank line
// in an .mtx body, an 'a' line of a .gr file that does not hold three
// numbers), the line is skipped or the read stops instead.
#include <sys/stat.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gi_internal.h"

namespace gi {
namespace {

struct EdgeList {
    std::vector<int32_t> u, v, w;
    void push(int32_t a, int32_t b, int32_t c) {
        u.push_back(a);
        v.push_back(b);
        w.push_back(c);
    }
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

// A cursor over text in memory.
struct Cursor {
    const char *p;
    const char *end;

    bool at_end() const { return p >= end; }
    int peek() const { return p < end ? (unsigned char)*p : -1; }
    void skip_space() {
        while (p < end && is_space(*p)) ++p;
    }
    // `in >> int32`: false (nothing but whitespace consumed) when no int32 is there.
    bool read_i32(int32_t *out) {
        skip_space();
        const char *q = p;
        bool neg = false;
        if (q < end && (*q == '+' || *q == '-')) {
            neg = *q == '-';
            ++q;
        }
        if (q >= end || *q < '0' || *q > '9') return false;
        int64_t x = 0;
        while (q < end && *q >= '0' && *q <= '9') {
            x = x * 10 + (*q - '0');
            if (x > (int64_t)INT32_MAX + 1) return false;  // out of range: the read fails
            ++q;
        }
        if (neg) x = -x;
        if (x < INT32_MIN || x > INT32_MAX) return false;
        *out = (int32_t)x;
        p = q;
        return true;
    }
    bool read_i64(int64_t *out) {
        skip_space();
        const char *q = p;
        bool neg = false;
        if (q < end && (*q == '+' || *q == '-')) {
            neg = *q == '-';
            ++q;
        }
        if (q >= end || *q < '0' || *q > '9') return false;
        int64_t x = 0;
        while (q < end && *q >= '0' && *q <= '9') {
            if (x > (INT64_MAX - 9) / 10) return false;
            x = x * 10 + (*q - '0');
            ++q;
        }
        *out = neg ? -x : x;
        p = q;
        return true;
    }
    // `in >> std::string`
    std::string read_word() {
        skip_space();
        const char *q = p;
        while (p < end && !is_space(*p)) ++p;
        return std::string(q, p);
    }
    // `in.ignore(200, '\n')`
    void ignore_line200() {
        for (int i = 0; i < 200 && p < end; ++i)
            if (*p++ == '\n') return;
    }
    // `std::getline(in, line)`: the line without its '\n'
    Cursor getline() {
        const char *q = p;
        while (p < end && *p != '\n') ++p;
        Cursor line{q, p};
        if (p < end) ++p;
        return line;
    }
};

bool ends_with(const std::string &s, const char *suffix) {
    const size_t n = strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

std::string suffix_of(const std::string &path) {
    const size_t dot = path.rfind('.');
    return dot == std::string::npos ? std::string() : path.substr(dot);
}

int read_whole(const std::string &path, std::vector<char> *buf) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return fail(GI_ERR_INVALID, "cannot open %s", path.c_str());
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return fail(GI_ERR_INVALID, "cannot seek %s", path.c_str());
    }
    const long n = ftell(f);
    rewind(f);
    buf->resize(n > 0 ? (size_t)n : 0);
    const size_t got = n > 0 ? fread(buf->data(), 1, (size_t)n, f) : 0;
    fclose(f);
    if (n < 0 || got != (size_t)n) return fail(GI_ERR_INVALID, "cannot read %s", path.c_str());
    return GI_OK;
}

// reader.h:51-68: pairs (weight 1) or triples, until the first read fails.
void parse_el(Cursor c, bool weighted, EdgeList *el) {
    int32_t a, b, w = 1;
    while (c.read_i32(&a) && c.read_i32(&b) && (!weighted || c.read_i32(&w))) el->push(a, b, w);
}

// reader.h:70-85: "a u v w" lines; every other line is skipped 200
// characters at a time, as `ignore(200, '\n')` does.
void parse_gr(Cursor c, EdgeList *el) {
    while (!c.at_end()) {
        if (c.peek() == 'a') {
            ++c.p;
            int32_t a, b, w;
            if (!(c.read_i32(&a) && c.read_i32(&b) && c.read_i32(&w))) return;
            el->push(a, b, w);
        } else {
            c.ignore_line200();
        }
    }
}

// reader.h:88-141. Ids from 1; every non-comment line after the header is
// the next vertex's adjacency, an empty one included; the last entry of a
// line is read only when whitespace follows it, as the reference reads.
int parse_metis(Cursor c, EdgeList *el) {
    int32_t n = 0, m = 0;
    bool weights = false;
    for (;;) {
        if (c.at_end()) return fail(GI_ERR_INVALID, "METIS file without a header");
        if (c.peek() == '%') {
            c.ignore_line200();
            continue;
        }
        Cursor h = c.getline();
        if (!h.read_i32(&n) || !h.read_i32(&m)) return fail(GI_ERR_INVALID, "bad METIS header");
        h.skip_space();
        if (!h.at_end()) {
            int32_t fmt = 0;
            if (!h.read_i32(&fmt)) return fail(GI_ERR_INVALID, "bad METIS fmt");
            if (fmt == 1) weights = true;
            else if (fmt != 0 && fmt != 100) return fail(GI_ERR_UNSUPPORTED, "METIS fmt %d is not supported", fmt);
        }
        break;
    }
    int32_t u = 0;
    while (u < n) {
        if (c.peek() == '%') {
            c.ignore_line200();
            continue;
        }
        Cursor line = c.getline();  // past the end: an empty line
        int32_t v, w = 1;
        // The reference's loop is `while (edge_stream >> v >> std::ws)`: a
        // value whose digits end the line leaves the stream at its end, the
        // std::ws read then fails and the value is not taken (libstdc++). So
        // the last entry of a line counts only when whitespace follows it.
        while (line.read_i32(&v) && (!weights || line.read_i32(&w)) && !line.at_end()) el->push(u, v - 1, w);
        ++u;
    }
    return GI_OK;
}

// reader.h:145-213.
int parse_mtx(Cursor c, EdgeList *el) {
    const std::string start = c.read_word(), object = c.read_word(), format = c.read_word(),
                      field = c.read_word(), symmetry = c.read_word();
    c.skip_space();
    if (start != "%%MatrixMarket") return fail(GI_ERR_INVALID, ".mtx file did not start with %%%%MatrixMarket");
    if (object != "matrix" || format != "coordinate")
        return fail(GI_ERR_UNSUPPORTED, "only matrix coordinate .mtx files are read");
    if (field == "complex") return fail(GI_ERR_UNSUPPORTED, "complex .mtx weights are not supported");
    bool weights;
    if (field == "pattern") weights = false;
    else if (field == "real" || field == "double" || field == "integer") weights = true;
    else return fail(GI_ERR_UNSUPPORTED, "unrecognized .mtx field type %s", field.c_str());
    bool undirected;
    if (symmetry == "symmetric") undirected = true;
    else if (symmetry == "general" || symmetry == "skew-symmetric") undirected = false;
    else return fail(GI_ERR_UNSUPPORTED, "unsupported .mtx symmetry %s", symmetry.c_str());
    while (c.peek() == '%') c.ignore_line200();
    int64_t rows, cols, nnz;
    if (!c.read_i64(&rows) || !c.read_i64(&cols) || !c.read_i64(&nnz)) return fail(GI_ERR_INVALID, "bad .mtx size line");
    c.skip_space();
    if (rows != cols) return fail(GI_ERR_INVALID, ".mtx matrix must be square (%lld x %lld)", (long long)rows, (long long)cols);
    while (!c.at_end()) {
        Cursor line = c.getline();
        int32_t a, b, w = 1;
        if (!line.read_i32(&a) || !line.read_i32(&b)) continue;  // blank line
        if (weights && !line.read_i32(&w)) w = 0;
        el->push(a - 1, b - 1, w);
        if (undirected) el->push(b - 1, a - 1, w);
    }
    return GI_OK;
}

// Order of (neighbour, weight) pairs as GAPBS's NodeWeight::operator<
// (signed weights): one unsigned 64-bit key.
inline uint64_t pair_key(int32_t v, int32_t w) {
    return ((uint64_t)(uint32_t)v << 32) | (uint64_t)((uint32_t)w ^ 0x80000000u);
}

// builder.h:62-71, 134-226: V = largest id + 1, lists sorted by (v, w),
// repeated neighbours kept once with the least weight, self-loops dropped.
int build_csr(const EdgeList &el, bool symmetrize, bool weighted, gi_host_csr *out) {
    const int64_t m = (int64_t)el.u.size();
    int64_t vmax = 0;
    for (int64_t i = 0; i < m; ++i) {
        if (el.u[i] < 0 || el.v[i] < 0) return fail(GI_ERR_RANGE, "negative vertex id in edge %lld", (long long)i);
        vmax = std::max<int64_t>(vmax, std::max(el.u[i], el.v[i]));
    }
    if (vmax >= INT32_MAX) return fail(GI_ERR_RANGE, "vertex ids exceed int32");
    const int32_t V = (int32_t)(vmax + 1);  // an empty list gives one vertex, as FindMaxNodeID does
    std::vector<int64_t> start((size_t)V + 1, 0);
    for (int64_t i = 0; i < m; ++i) {
        ++start[(size_t)el.u[i] + 1];
        if (symmetrize) ++start[(size_t)el.v[i] + 1];
    }
    for (int32_t v = 0; v < V; ++v) start[(size_t)v + 1] += start[(size_t)v];
    std::vector<uint64_t> keys((size_t)start[V]);
    {
        std::vector<int64_t> pos(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < m; ++i) {
            keys[(size_t)pos[(size_t)el.u[i]]++] = pair_key(el.v[i], el.w[i]);
            if (symmetrize) keys[(size_t)pos[(size_t)el.v[i]]++] = pair_key(el.u[i], el.w[i]);
        }
    }
    // per list: sort, keep the first of each neighbour, drop the self-loop;
    // lists split over host threads by edge count
    std::vector<int32_t> kept((size_t)V, 0);
    auto squish = [&](int32_t v0, int32_t v1) {
        for (int32_t v = v0; v < v1; ++v) {
            uint64_t *b = keys.data() + start[v], *e = keys.data() + start[(size_t)v + 1];
            std::sort(b, e);
            int64_t n = 0;
            for (uint64_t *k = b; k < e; ++k) {
                const int32_t t = (int32_t)(*k >> 32);
                if (t == v) continue;
                if (n > 0 && (int32_t)(b[n - 1] >> 32) == t) continue;
                b[n++] = *k;
            }
            kept[(size_t)v] = (int32_t)n;
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int64_t per = start[V] / hw + 1;
    std::vector<std::thread> pool;
    int32_t v0 = 0;
    for (unsigned t = 0; t < hw && v0 < V; ++t) {
        int32_t v1 = v0;
        while (v1 < V && start[(size_t)v1] - start[(size_t)v0] < per) ++v1;
        if (t == hw - 1) v1 = V;
        try {
            pool.emplace_back(squish, v0, v1);
        } catch (...) {  // no thread to be had: this range runs here
            squish(v0, v1);
        }
        v0 = v1;
    }
    if (v0 < V) squish(v0, V);
    for (auto &th : pool) th.join();
    int64_t E = 0;
    for (int32_t v = 0; v < V; ++v) E += kept[(size_t)v];
    if (E > INT32_MAX) return fail(GI_ERR_RANGE, "more than 2^31 edges");
    out->num_vertices = V;
    out->num_edges = E;
    out->directed = symmetrize ? 0 : 1;
    out->offsets = (int32_t *)malloc(sizeof(int32_t) * ((size_t)V + 1));
    out->nbrs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    out->weights = weighted ? (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)) : nullptr;
    if (!out->offsets || !out->nbrs || (weighted && !out->weights)) return fail(GI_ERR_OOM, "host allocation failed");
    int64_t o = 0;
    for (int32_t v = 0; v < V; ++v) {
        out->offsets[v] = (int32_t)o;
        const uint64_t *b = keys.data() + start[v];
        for (int32_t i = 0; i < kept[(size_t)v]; ++i, ++o) {
            out->nbrs[o] = (int32_t)(b[i] >> 32);
            if (weighted) out->weights[o] = (int32_t)((uint32_t)b[i] ^ 0x80000000u);
        }
    }
    out->offsets[V] = (int32_t)o;
    return GI_OK;
}

int check_csr(const gi_host_csr &c) {
    if (c.offsets[0] != 0 || c.offsets[c.num_vertices] != c.num_edges) return fail(GI_ERR_INVALID, "bad offsets");
    for (int32_t v = 0; v < c.num_vertices; ++v)
        if (c.offsets[v + 1] < c.offsets[v]) return fail(GI_ERR_INVALID, "offsets not monotone at %d", v);
    for (int64_t e = 0; e < c.num_edges; ++e)
        if (c.nbrs[e] < 0 || c.nbrs[e] >= c.num_vertices) return fail(GI_ERR_RANGE, "neighbour out of range at %lld", (long long)e);
    return GI_OK;
}

// reader.h:332-373 (the inverse CSR of a directed graph is not needed: the
// device builds transposes itself).
int read_serialized(const std::vector<char> &buf, bool weighted, gi_host_csr *out) {
    size_t at = 0;
    auto take = [&](void *dst, size_t n) {
        if (at + n > buf.size()) return false;
        memcpy(dst, buf.data() + at, n);
        at += n;
        return true;
    };
    uint8_t directed = 0;
    int64_t E = 0, V = 0;
    if (!take(&directed, 1) || !take(&E, 8) || !take(&V, 8)) return fail(GI_ERR_INVALID, "truncated serialized graph");
    if (V < 0 || V > INT32_MAX - 1 || E < 0 || E > INT32_MAX) return fail(GI_ERR_RANGE, "serialized graph too large");
    // the header's sizes are checked against the bytes present before any
    // allocation: a truncated or corrupt header must not request gigabytes
    const uint64_t need = 8ull * ((uint64_t)V + 1) + (weighted ? 8ull : 4ull) * (uint64_t)E;
    if (need > (uint64_t)(buf.size() - at)) return fail(GI_ERR_INVALID, "truncated serialized graph");
    std::vector<int64_t> off((size_t)V + 1);
    if (!take(off.data(), sizeof(int64_t) * ((size_t)V + 1))) return fail(GI_ERR_INVALID, "truncated offsets");
    out->num_vertices = (int32_t)V;
    out->num_edges = E;
    out->directed = directed ? 1 : 0;
    out->offsets = (int32_t *)malloc(sizeof(int32_t) * ((size_t)V + 1));
    out->nbrs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    out->weights = weighted ? (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1)) : nullptr;
    if (!out->offsets || !out->nbrs || (weighted && !out->weights)) return fail(GI_ERR_OOM, "host allocation failed");
    for (int64_t v = 0; v <= V; ++v) {
        if (off[(size_t)v] < 0 || off[(size_t)v] > E) return fail(GI_ERR_INVALID, "bad serialized offsets");
        out->offsets[v] = (int32_t)off[(size_t)v];
    }
    if (weighted) {
        std::vector<int32_t> pairs((size_t)E * 2);
        if (!take(pairs.data(), sizeof(int32_t) * 2 * (size_t)E)) return fail(GI_ERR_INVALID, "truncated neighbours");
        for (int64_t e = 0; e < E; ++e) {
            out->nbrs[e] = pairs[(size_t)(2 * e)];
            out->weights[e] = pairs[(size_t)(2 * e + 1)];
        }
    } else if (!take(out->nbrs, sizeof(int32_t) * (size_t)E)) {
        return fail(GI_ERR_INVALID, "truncated neighbours");
    }
    return check_csr(*out);
}

// graph.h:252-268: int32 V, int32 E, src[E], dst[E], weight[E], offsets[V+1].
int read_graphit_bin(const std::vector<char> &buf, gi_host_csr *out) {
    int32_t hdr[2];
    if (buf.size() < sizeof(hdr)) return fail(GI_ERR_INVALID, "truncated .graphit_bin");
    memcpy(hdr, buf.data(), sizeof(hdr));
    const int32_t V = hdr[0], E = hdr[1];
    if (V < 0 || E < 0) return fail(GI_ERR_INVALID, "bad .graphit_bin header");
    const size_t need = sizeof(hdr) + sizeof(int32_t) * (3 * (size_t)E + (size_t)V + 1);
    if (buf.size() < need) return fail(GI_ERR_INVALID, "truncated .graphit_bin");
    const int32_t *body = reinterpret_cast<const int32_t *>(buf.data() + sizeof(hdr));
    out->num_vertices = V;
    out->num_edges = E;
    out->directed = 1;
    out->offsets = (int32_t *)malloc(sizeof(int32_t) * ((size_t)V + 1));
    out->nbrs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    out->weights = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    if (!out->offsets || !out->nbrs || !out->weights) return fail(GI_ERR_OOM, "host allocation failed");
    memcpy(out->nbrs, body + E, sizeof(int32_t) * (size_t)E);
    memcpy(out->weights, body + 2 * (size_t)E, sizeof(int32_t) * (size_t)E);
    memcpy(out->offsets, body + 3 * (size_t)E, sizeof(int32_t) * ((size_t)V + 1));
    return check_csr(*out);
}

int read_file(const std::string &path, int32_t flags, gi_host_csr *out) {
    memset(out, 0, sizeof(*out));
    const bool sym = (flags & GI_LOAD_SYMMETRIZE) != 0;
    const std::string suf = suffix_of(path);
    std::vector<char> buf;
    GI_TRY(read_whole(path, &buf));
    if (ends_with(path, ".graphit_bin")) return read_graphit_bin(buf, out);
    if (suf == ".sg" || suf == ".wsg") return read_serialized(buf, suf == ".wsg", out);
    Cursor c{buf.data(), buf.data() + buf.size()};
    EdgeList el;
    if (suf == ".el") parse_el(c, false, &el);
    else if (suf == ".wel") parse_el(c, true, &el);
    else if (suf == ".gr") parse_gr(c, &el);
    else if (suf == ".graph") GI_TRY(parse_metis(c, &el));
    else if (suf == ".mtx") GI_TRY(parse_mtx(c, &el));
    else return fail(GI_ERR_UNSUPPORTED, "unrecognized graph file suffix '%s'", suf.c_str());
    return build_csr(el, sym, true, out);  // every text format carries weights (1 when absent)
}

}  // namespace
}  // namespace gi

using namespace gi;

extern "C" {

void gi_host_csr_free(gi_host_csr *c) {
    if (!c) return;
    free(c->offsets);
    free(c->nbrs);
    free(c->weights);
    memset(c, 0, sizeof(*c));
}

int gi_read_graph_file(const char *path, int32_t flags, gi_host_csr *out) {
    if (!path || !out) return fail(GI_ERR_INVALID, "null argument");
    const int rc = read_file(path, flags, out);
    if (rc != GI_OK) gi_host_csr_free(out);
    return rc;
}

int gi_write_graphit_bin(const gi_host_csr *c, const char *path) {
    if (!c || !path || !c->offsets || (c->num_edges > 0 && !c->nbrs)) return fail(GI_ERR_INVALID, "null argument");
    if (c->num_edges > INT32_MAX) return fail(GI_ERR_RANGE, "more than 2^31 edges");
    FILE *f = fopen(path, "wb");
    if (!f) return fail(GI_ERR_INVALID, "cannot write %s", path);
    const int32_t V = c->num_vertices, E = (int32_t)c->num_edges;
    std::vector<int32_t> src((size_t)E), w((size_t)E, 1);
    for (int32_t v = 0; v < V; ++v)
        for (int32_t e = c->offsets[v]; e < c->offsets[v + 1]; ++e) src[(size_t)e] = v;
    if (c->weights) memcpy(w.data(), c->weights, sizeof(int32_t) * (size_t)E);
    bool ok = fwrite(&V, sizeof(int32_t), 1, f) == 1 && fwrite(&E, sizeof(int32_t), 1, f) == 1 &&
              fwrite(src.data(), sizeof(int32_t), (size_t)E, f) == (size_t)E &&
              fwrite(c->nbrs, sizeof(int32_t), (size_t)E, f) == (size_t)E &&
              fwrite(w.data(), sizeof(int32_t), (size_t)E, f) == (size_t)E &&
              fwrite(c->offsets, sizeof(int32_t), (size_t)V + 1, f) == (size_t)V + 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? GI_OK : fail(GI_ERR_INVALID, "short write to %s", path);
}

int gi_graph_load(const char *path, int32_t flags, gi_graph **out) {
    if (!path || !out) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    const std::string p(path);
    const bool sym = (flags & GI_LOAD_SYMMETRIZE) != 0;
    gi_host_csr c;
    memset(&c, 0, sizeof(c));
    int rc;
    if ((flags & GI_LOAD_CACHE) && !ends_with(p, ".graphit_bin")) {
        // graph.h:239-248: "<file>.<type id of int32>.graphit_bin" next to the file
        const std::string cache = p + (sym ? ".sym" : "") + ".i.graphit_bin";
        struct stat sb;
        if (stat(cache.c_str(), &sb) == 0) {
            rc = gi_read_graph_file(cache.c_str(), 0, &c);
            c.directed = sym ? 0 : 1;
        } else {
            rc = gi_read_graph_file(path, flags, &c);
            if (rc == GI_OK) (void)gi_write_graphit_bin(&c, cache.c_str());  // a cache it cannot write is skipped
        }
    } else {
        rc = gi_read_graph_file(path, flags, &c);
    }
    if (rc == GI_OK)
        rc = gi_graph_from_csr(c.num_vertices, c.num_edges, c.offsets, c.nbrs, c.weights, c.directed, out);
    gi_host_csr_free(&c);
    return rc;
}

}  // extern "C"
