// The reference's on-disk index format: Go's encoding/gob stream of
// serializedPQIVF (pqivf/index.go:968-1101).
//
//   GobEncode / GobDecode  one gob stream: the type definitions the value
//                          needs, then the serializedPQIVF value
//   Save / Load            an outer gob stream holding one value of the
//                          *Index GobEncoder type: that stream's bytes
//
// Wire format as documented by Go's encoding/gob: messages are a byte count,
// a signed type id (negative: a type definition follows) and the payload.
// Unsigned integers below 128 take one byte, larger ones a byte holding the
// negated byte count and then the big-endian bytes without leading zeros;
// signed integers are zig-zag folded (bit 0 holds the sign); floats travel
// as float64 bits, byte-reversed, then as an unsigned integer; strings and
// byte slices are a count and the bytes. A struct is a list of (field-number
// delta, value) pairs ending with a 0 delta, omitting zero fields; slices
// and maps send a count and every element (empty ones inside a struct are
// omitted, except non-nil maps, which Go always sends). A top-level value of
// a non-struct type is preceded by a 0 delta.
//
// The writer emits the type definitions Go's encoder emits for this struct,
// in its order, with its type ids and names, and visits maps in ascending
// key order (Go's map order is random); so a file in the layout before the
// FormatVersion field existed re-encodes byte for byte. The reader matches
// struct fields by name, skips fields it does not know, and checks each
// known field's wire type, as Go's decoder does.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "pqivf.h"

namespace hann {
namespace {

// predefined gob type ids
constexpr int T_BOOL = 1, T_INT = 2, T_UINT = 3, T_FLOAT = 4, T_BYTES = 5, T_STRING = 6, T_COMPLEX = 7,
              T_INTERFACE = 8;

// the type ids Go's encoder gives serializedPQIVF and its parts in a fresh
// stream (as in the reference's own saved index)
constexpr int ID_INDEX_PTR = 64, ID_INDEX = 65, ID_SER = 66, ID_F32S = 67, ID_F32SS = 68, ID_MAP_II = 69,
              ID_ENTRY = 70, ID_INTS = 71, ID_ENTRIES = 72, ID_MAP_LISTS = 73, ID_F32SSS = 74, ID_MAP_PEND = 75;

constexpr int kGobFormatVersion = 1;   // formatVersion (pqivf/index.go:986-989)

[[noreturn]] void bad(const std::string &what) { fail(HANN_ERR_FORMAT, "gob: " + what); }

// ------------------------------------------------------------- writing --
struct Out {
    std::vector<unsigned char> b;
    void u(uint64_t x) {
        if (x < 128) {
            b.push_back((unsigned char)x);
            return;
        }
        unsigned char t[8];
        int n = 0;
        while (x) {
            t[n++] = (unsigned char)(x & 0xff);
            x >>= 8;
        }
        b.push_back((unsigned char)(256 - n));
        while (n) b.push_back(t[--n]);
    }
    void i(int64_t v) { u(v < 0 ? ((uint64_t)~v << 1) | 1u : (uint64_t)v << 1); }
    void f(float v) {
        const double d = (double)v;
        uint64_t bits;
        std::memcpy(&bits, &d, 8);
        u(__builtin_bswap64(bits));
    }
    void s(const std::string &v) {
        u(v.size());
        b.insert(b.end(), v.begin(), v.end());
    }
    void bytes(const std::vector<unsigned char> &v) {
        u(v.size());
        b.insert(b.end(), v.begin(), v.end());
    }
    // one message: the count, then the type id and payload held in p
    void message(const Out &p) {
        u(p.b.size());
        b.insert(b.end(), p.b.begin(), p.b.end());
    }
};

// struct fields by index: the delta from the previous sent field
struct Fields {
    Out &o;
    int last = -1;
    void at(int idx) {
        o.u((uint64_t)(idx - last));
        last = idx;
    }
    void end() { o.u(0); }
};

void put_common(Out &p, const std::string &name, int id) {
    Fields f{p};
    f.at(0);   // CommonType
    Fields c{p};
    if (!name.empty()) {
        c.at(0);
        p.s(name);
    }
    c.at(1);
    p.i(id);
    c.end();
}

// wireType fields: ArrayT 0, SliceT 1, StructT 2, MapT 3, GobEncoderT 4
void def_slice(Out &o, int id, const std::string &name, int elem) {
    Out p;
    p.i(-id);
    Fields w{p};
    w.at(1);
    Fields t{p};
    put_common(p, name, id);
    t.last = 0;
    t.at(1);
    p.i(elem);
    t.end();
    w.end();
    o.message(p);
}

void def_map(Out &o, int id, const std::string &name, int key, int elem) {
    Out p;
    p.i(-id);
    Fields w{p};
    w.at(3);
    Fields t{p};
    put_common(p, name, id);
    t.last = 0;
    t.at(1);
    p.i(key);
    t.at(2);
    p.i(elem);
    t.end();
    w.end();
    o.message(p);
}

void def_struct(Out &o, int id, const std::string &name, const std::vector<std::pair<std::string, int>> &fields) {
    Out p;
    p.i(-id);
    Fields w{p};
    w.at(2);
    Fields t{p};
    put_common(p, name, id);
    t.last = 0;
    if (!fields.empty()) {
        t.at(1);
        p.u(fields.size());
        for (const auto &fd : fields) {
            Fields ft{p};
            if (!fd.first.empty()) {
                ft.at(0);
                p.s(fd.first);
            }
            ft.at(1);
            p.i(fd.second);
            ft.end();
        }
    }
    t.end();
    w.end();
    o.message(p);
}

void def_gob_encoder(Out &o, int id, int common_id) {
    Out p;
    p.i(-id);
    Fields w{p};
    w.at(4);
    Fields t{p};
    put_common(p, "", common_id);
    t.end();
    w.end();
    o.message(p);
}

void put_floats(Out &p, const float *v, int64_t n) {
    p.u((uint64_t)n);
    for (int64_t j = 0; j < n; ++j) p.f(v[j]);
}

// ------------------------------------------------------------- reading --
struct In {
    const unsigned char *p, *e;
    bool done() const { return p >= e; }
    uint64_t u() {
        if (p >= e) bad("unexpected end of data");
        const unsigned c = *p++;
        if (c < 128) return c;
        const int n = 256 - (int)c;
        if (n > 8) bad("invalid unsigned integer");
        if (e - p < n) bad("unexpected end of data");
        uint64_t x = 0;
        for (int j = 0; j < n; ++j) x = (x << 8) | *p++;
        return x;
    }
    int64_t i() {
        const uint64_t x = u();
        return (x & 1) ? ~(int64_t)(x >> 1) : (int64_t)(x >> 1);
    }
    double f() {
        const uint64_t x = __builtin_bswap64(u());
        double d;
        std::memcpy(&d, &x, 8);
        return d;
    }
    // a count of elements that each take at least one byte
    size_t count() {
        const uint64_t n = u();
        if (n > (uint64_t)(e - p)) bad("element count exceeds the data");
        return (size_t)n;
    }
    In take(size_t n) {
        if (n > (size_t)(e - p)) bad("unexpected end of data");
        In r{p, p + n};
        p += n;
        return r;
    }
    std::string str() {
        const size_t n = count();
        std::string s(reinterpret_cast<const char *>(p), n);
        p += n;
        return s;
    }
};

enum Kind { K_NONE, K_STRUCT, K_SLICE, K_ARRAY, K_MAP, K_GOBENC };

struct WType {
    Kind kind = K_NONE;
    std::vector<std::pair<std::string, int>> fields;
    int elem = 0, key = 0;
};

struct Stream {
    In r;
    std::unordered_map<int, WType> types;

    // struct walk over the bootstrap types: calls fn(field index) per sent field
    template <typename F>
    static void walk(In &m, F &&fn) {
        int f = -1;
        for (;;) {
            const uint64_t d = m.u();
            if (d == 0) return;
            if (d > 64) bad("invalid field delta");
            f += (int)d;
            fn(f);
        }
    }
    static void common(In &m) {
        walk(m, [&](int f) {
            if (f == 0) m.str();
            else if (f == 1) m.i();
            else bad("invalid type definition");
        });
    }
    WType wire(In &m) {
        WType t;
        int seen = 0;
        walk(m, [&](int f) {
            ++seen;
            switch (f) {
            case 0:   // arrayType {CommonType, Elem, Len}
                t.kind = K_ARRAY;
                walk(m, [&](int g) {
                    if (g == 0) common(m);
                    else if (g == 1) t.elem = (int)m.i();
                    else if (g == 2) m.i();
                    else bad("invalid array type");
                });
                break;
            case 1:   // sliceType {CommonType, Elem}
                t.kind = K_SLICE;
                walk(m, [&](int g) {
                    if (g == 0) common(m);
                    else if (g == 1) t.elem = (int)m.i();
                    else bad("invalid slice type");
                });
                break;
            case 2:   // structType {CommonType, Field []fieldType}
                t.kind = K_STRUCT;
                walk(m, [&](int g) {
                    if (g == 0) common(m);
                    else if (g == 1) {
                        const size_t n = m.count();
                        for (size_t j = 0; j < n; ++j) {
                            std::pair<std::string, int> fd{"", 0};
                            walk(m, [&](int h) {
                                if (h == 0) fd.first = m.str();
                                else if (h == 1) fd.second = (int)m.i();
                                else bad("invalid struct field");
                            });
                            t.fields.push_back(fd);
                        }
                    } else bad("invalid struct type");
                });
                break;
            case 3:   // mapType {CommonType, Key, Elem}
                t.kind = K_MAP;
                walk(m, [&](int g) {
                    if (g == 0) common(m);
                    else if (g == 1) t.key = (int)m.i();
                    else if (g == 2) t.elem = (int)m.i();
                    else bad("invalid map type");
                });
                break;
            case 4: case 5: case 6:   // GobEncoder, BinaryMarshaler, TextMarshaler
                t.kind = K_GOBENC;
                walk(m, [&](int g) {
                    if (g == 0) common(m);
                    else bad("invalid encoder type");
                });
                break;
            default:
                bad("invalid type definition");
            }
        });
        if (seen != 1) bad("invalid type definition");
        return t;
    }
    // the next value message: its type id, the payload in body
    int next(In &body) {
        for (;;) {
            if (r.done()) bad("no value in the stream");
            const size_t n = (size_t)r.u();
            In m = r.take(n);
            const int64_t id = m.i();
            if (id < 0) {
                if (-id > ((int64_t)1 << 30) || -id < 64) bad("invalid type id in a definition");
                types[(int)-id] = wire(m);
                continue;
            }
            if (id == 0) bad("invalid type id");
            body = m;
            return (int)id;
        }
    }
    const WType &type(int id) const {
        auto it = types.find(id);
        if (it == types.end()) bad("reference to an undefined type " + std::to_string(id));
        return it->second;
    }
    void skip(In &m, int id, int depth = 0) const {
        if (depth > 64) bad("values nested too deep");
        switch (id) {
        case T_BOOL: case T_INT: case T_UINT: case T_FLOAT: m.u(); return;
        case T_BYTES: case T_STRING: m.str(); return;
        case T_COMPLEX: m.u(); m.u(); return;
        case T_INTERFACE: bad("interface values are not supported");
        default: break;
        }
        const WType &t = type(id);
        switch (t.kind) {
        case K_STRUCT:
            walk(m, [&](int f) {
                if (f >= (int)t.fields.size()) bad("field number out of range");
                skip(m, t.fields[f].second, depth + 1);
            });
            return;
        case K_SLICE: case K_ARRAY: {
            const size_t n = m.count();
            for (size_t j = 0; j < n; ++j) skip(m, t.elem, depth + 1);
            return;
        }
        case K_MAP: {
            const size_t n = m.count();
            for (size_t j = 0; j < n; ++j) {
                skip(m, t.key, depth + 1);
                skip(m, t.elem, depth + 1);
            }
            return;
        }
        case K_GOBENC: m.str(); return;
        default: bad("invalid type");
        }
    }
    // typed readers: the wire type must match the field's Go type
    static void expect(bool ok, const char *field) {
        if (!ok) bad(std::string("type mismatch for field ") + field);
    }
    bool is_list_of(int id, int elem) const {
        if (id < 64) return false;
        const WType &t = type(id);
        return (t.kind == K_SLICE || t.kind == K_ARRAY) && t.elem == elem;
    }
    static float f32(double v) {
        if (std::isfinite(v) && std::fabs(v) > (double)FLT_MAX) bad("float32 value out of range");
        return (float)v;
    }
    std::vector<float> floats(In &m, int id, const char *field) const {
        expect(is_list_of(id, T_FLOAT), field);
        const size_t n = m.count();
        std::vector<float> v(n);
        for (size_t j = 0; j < n; ++j) v[j] = f32(m.f());
        return v;
    }
    std::vector<std::vector<float>> floats2(In &m, int id, const char *field) const {
        expect(id >= 64 && (type(id).kind == K_SLICE || type(id).kind == K_ARRAY), field);
        const int e = type(id).elem;
        const size_t n = m.count();
        std::vector<std::vector<float>> v(n);
        for (size_t j = 0; j < n; ++j) v[j] = floats(m, e, field);
        return v;
    }
};

struct GobEntry {
    int64_t id = 0;
    std::vector<float> vec;
    std::vector<int64_t> codes;
};

}  // namespace

// GobEncode (pqivf/index.go:991-1017): format_version 1 writes the current
// field set; 0 writes the layout from before FormatVersion and
// NumCandidateClusters existed.
std::vector<unsigned char> PQIVF::gob_encode(int format_version) const {
    if (format_version != 0 && format_version != kGobFormatVersion)
        fail(HANN_ERR_INVALID_ARG, strf("gob format version %d is not supported (0 or %d)", format_version,
                                        kGobFormatVersion));
    Out o;
    std::vector<std::pair<std::string, int>> fields = {
        {"Dimension", T_INT}, {"CoarseK", T_INT}, {"CoarseCentroids", ID_F32SS}, {"ClusterCounts", ID_MAP_II},
        {"InvertedLists", ID_MAP_LISTS}, {"NumSubquantizers", T_INT}, {"Codebooks", ID_F32SSS}, {"PqK", T_INT},
        {"KMeansIters", T_INT}, {"AllowBruteForceFallback", T_BOOL}, {"Trained", T_BOOL},
        {"PendingVectors", ID_MAP_PEND}};
    if (format_version > 0) {
        fields.push_back({"FormatVersion", T_INT});
        fields.push_back({"NumCandidateClusters", T_INT});
    }
    def_struct(o, ID_SER, "serializedPQIVF", fields);
    def_slice(o, ID_F32SS, "[][]float32", ID_F32S);
    def_slice(o, ID_F32S, "", T_FLOAT);
    def_map(o, ID_MAP_II, "map[int]int", T_INT, T_INT);
    def_map(o, ID_MAP_LISTS, "map[int][]pqivf.pqEntry", T_INT, ID_ENTRIES);
    def_slice(o, ID_ENTRIES, "", ID_ENTRY);
    def_struct(o, ID_ENTRY, "pqEntry", {{"ID", T_INT}, {"Vector", ID_F32S}, {"Codes", ID_INTS}, {"Cluster", T_INT}});
    def_slice(o, ID_INTS, "[]int", T_INT);
    def_slice(o, ID_F32SSS, "[][][]float32", ID_F32SS);
    def_map(o, ID_MAP_PEND, "map[int][]float32", T_INT, ID_F32S);

    Out p;
    p.i(ID_SER);
    Fields f{p};
    if (dim) { f.at(0); p.i(dim); }
    if (coarse_k) { f.at(1); p.i(coarse_k); }
    if (kc > 0) {
        f.at(2);
        p.u((uint64_t)kc);
        for (int64_t c = 0; c < kc; ++c) put_floats(p, &coarse[(size_t)(c * dim)], dim);
    }
    // ClusterCounts and InvertedLists share their keys: a list's count is its
    // length (pqivf/index.go:258,334,375,503)
    std::vector<int64_t> keys;
    for (size_t c = 0; c < lists.size(); ++c)
        if (lists[c].exists) keys.push_back((int64_t)c);
    f.at(3);
    p.u(keys.size());
    for (int64_t c : keys) {
        p.i(c);
        p.i((int64_t)lists[(size_t)c].ids.size());
    }
    f.at(4);
    p.u(keys.size());
    for (int64_t c : keys) {
        const ListData &L = lists[(size_t)c];
        p.i(c);
        p.u(L.ids.size());
        for (size_t r = 0; r < L.ids.size(); ++r) {
            Fields e{p};
            if (L.ids[r]) { e.at(0); p.i(L.ids[r]); }
            if (dim) { e.at(1); put_floats(p, &L.vecs[r * (size_t)dim], dim); }
            if (L.has_codes[r] && m > 0) {
                e.at(2);
                p.u((uint64_t)m);
                for (int64_t j = 0; j < m; ++j) p.i(L.codes[r * (size_t)m + (size_t)j]);
            }
            if (c) { e.at(3); p.i(c); }
            e.end();
        }
    }
    if (m) { f.at(5); p.i(m); }
    if (have_codebooks && m > 0) {
        const int64_t s = dim / m;
        f.at(6);
        p.u((uint64_t)m);
        for (int64_t i = 0; i < m; ++i) {
            p.u((uint64_t)cb_len[(size_t)i]);
            for (int64_t j = 0; j < cb_len[(size_t)i]; ++j)
                put_floats(p, &codebooks[(size_t)((i * pmax + j) * s)], s);
        }
    }
    if (pq_k) { f.at(7); p.i(pq_k); }
    if (iters) { f.at(8); p.i(iters); }
    if (fallback) { f.at(9); p.u(1); }
    if (trained) { f.at(10); p.u(1); }
    f.at(11);
    std::vector<std::pair<int64_t, size_t>> pend;
    for (size_t j = 0; j < pending.size(); ++j) pend.push_back({pending.ids[j], j});
    std::sort(pend.begin(), pend.end());
    p.u(pend.size());
    for (const auto &pr : pend) {
        p.i(pr.first);
        put_floats(p, &pending.vecs[pr.second * (size_t)dim], dim);
    }
    if (format_version > 0) {
        f.at(12);
        p.i(kGobFormatVersion);
        if (nprobe) { f.at(13); p.i(nprobe); }
    }
    f.end();
    o.message(p);
    return std::move(o.b);
}

// Save (pqivf/index.go:1086-1092): one value of the *Index GobEncoder type
std::vector<unsigned char> PQIVF::gob_save(int format_version) const {
    const std::vector<unsigned char> inner = gob_encode(format_version);
    Out o;
    def_gob_encoder(o, ID_INDEX_PTR, ID_INDEX);
    Out p;
    p.i(ID_INDEX_PTR);
    p.u(0);   // a non-struct value is sent as field 0 of a singleton
    p.bytes(inner);
    o.message(p);
    return std::move(o.b);
}

// GobDecode (pqivf/index.go:1020-1084): decode, check the version and the
// parameters with the reference's messages, then replace the state. The
// device index also requires what the reference assumes of a real file:
// centroids and pending vectors of dimension values, codebooks of dim/m
// columns, entries of dimension values with no or numSubquantizers codes
// inside their codebooks, and list keys naming a centroid. A list's count
// and its entries' Cluster field follow from its key.
void PQIVF::gob_decode(const void *buf, size_t len) {
    const unsigned char *b = reinterpret_cast<const unsigned char *>(buf);
    Stream st{In{b, b + len}, {}};
    In m{nullptr, nullptr};
    const int id = st.next(m);
    if (id < 64 || st.type(id).kind != K_STRUCT) bad("the value is not a struct");
    const WType &t = st.type(id);
    int64_t ddim = 0, dck = 0, dm = 0, dpk = 0, dit = 0, dver = 0, dnc = 0;
    bool dfb = false, dtr = false;
    std::vector<std::vector<float>> C;
    std::vector<std::vector<std::vector<float>>> CB;
    bool have_cb = false;
    std::vector<std::pair<int64_t, std::vector<GobEntry>>> lists_in;
    std::vector<std::pair<int64_t, std::vector<float>>> pend_in;
    auto int_field = [&](int wid, const char *name) {
        Stream::expect(wid == T_INT, name);
        return m.i();
    };
    auto bool_field = [&](int wid, const char *name) {
        Stream::expect(wid == T_BOOL, name);
        return m.u() != 0;
    };
    Stream::walk(m, [&](int f) {
        if (f >= (int)t.fields.size()) bad("field number out of range");
        const std::string &name = t.fields[(size_t)f].first;
        const int wid = t.fields[(size_t)f].second;
        if (name == "Dimension") ddim = int_field(wid, "Dimension");
        else if (name == "CoarseK") dck = int_field(wid, "CoarseK");
        else if (name == "NumSubquantizers") dm = int_field(wid, "NumSubquantizers");
        else if (name == "PqK") dpk = int_field(wid, "PqK");
        else if (name == "KMeansIters") dit = int_field(wid, "KMeansIters");
        else if (name == "FormatVersion") dver = int_field(wid, "FormatVersion");
        else if (name == "NumCandidateClusters") dnc = int_field(wid, "NumCandidateClusters");
        else if (name == "AllowBruteForceFallback") dfb = bool_field(wid, "AllowBruteForceFallback");
        else if (name == "Trained") dtr = bool_field(wid, "Trained");
        else if (name == "CoarseCentroids") C = st.floats2(m, wid, "CoarseCentroids");
        else if (name == "Codebooks") {
            Stream::expect(wid >= 64 && st.type(wid).kind == K_SLICE, "Codebooks");
            const int e = st.type(wid).elem;
            const size_t n = m.count();
            CB.resize(n);
            for (size_t j = 0; j < n; ++j) CB[j] = st.floats2(m, e, "Codebooks");
            have_cb = n > 0;
        } else if (name == "InvertedLists") {
            Stream::expect(wid >= 64 && st.type(wid).kind == K_MAP && st.type(wid).key == T_INT, "InvertedLists");
            const int le = st.type(wid).elem;
            Stream::expect(le >= 64 && st.type(le).kind == K_SLICE, "InvertedLists");
            const int ee = st.type(le).elem;
            Stream::expect(ee >= 64 && st.type(ee).kind == K_STRUCT, "InvertedLists");
            const WType &et = st.type(ee);
            const size_t n = m.count();
            for (size_t j = 0; j < n; ++j) {
                const int64_t key = m.i();
                const size_t ne = m.count();
                std::vector<GobEntry> entries(ne);
                for (size_t r = 0; r < ne; ++r) {
                    GobEntry &E = entries[r];
                    Stream::walk(m, [&](int g) {
                        if (g >= (int)et.fields.size()) bad("field number out of range");
                        const std::string &fn = et.fields[(size_t)g].first;
                        const int fid = et.fields[(size_t)g].second;
                        if (fn == "ID") E.id = int_field(fid, "ID");
                        else if (fn == "Vector") E.vec = st.floats(m, fid, "Vector");
                        else if (fn == "Codes") {
                            Stream::expect(st.is_list_of(fid, T_INT), "Codes");
                            const size_t nc = m.count();
                            E.codes.resize(nc);
                            for (size_t q = 0; q < nc; ++q) E.codes[q] = m.i();
                        } else if (fn == "Cluster") int_field(fid, "Cluster");
                        else st.skip(m, fid);
                    });
                }
                lists_in.emplace_back(key, std::move(entries));
            }
        } else if (name == "PendingVectors") {
            Stream::expect(wid >= 64 && st.type(wid).kind == K_MAP && st.type(wid).key == T_INT, "PendingVectors");
            const int ve = st.type(wid).elem;
            const size_t n = m.count();
            for (size_t j = 0; j < n; ++j) {
                const int64_t key = m.i();
                pend_in.emplace_back(key, st.floats(m, ve, "PendingVectors"));
            }
        } else if (name == "ClusterCounts") {
            Stream::expect(wid >= 64 && st.type(wid).kind == K_MAP, "ClusterCounts");
            st.skip(m, wid);
        } else
            st.skip(m, wid);
    });
    if (dver > kGobFormatVersion)
        fail(HANN_ERR_FORMAT, strf("index file has format version %lld, but this build supports up to %d",
                                   (long long)dver, kGobFormatVersion));
    if (ddim <= 0) fail(HANN_ERR_FORMAT, strf("serialized index has invalid dimension %lld", (long long)ddim));
    if (dm <= 0 || ddim % dm != 0)
        fail(HANN_ERR_FORMAT, strf("serialized index has invalid subquantizer count %lld for dimension %lld",
                                   (long long)dm, (long long)ddim));
    if (dck <= 0 || dpk <= 0 || dit <= 0)
        fail(HANN_ERR_FORMAT,
             strf("serialized index has invalid training parameters: coarseK %lld, pqK %lld, kMeansIters %lld",
                  (long long)dck, (long long)dpk, (long long)dit));
    const int64_t s = ddim / dm;
    IndexState S;
    S.dim = ddim;
    S.coarse_k = dck;
    S.m = dm;
    S.pq_k = dpk;
    S.iters = dit;
    S.nprobe = dnc;
    S.fallback = dfb;
    S.trained = dtr;
    S.kc = (int64_t)C.size();
    S.coarse.reserve(C.size() * (size_t)ddim);
    for (const auto &row : C) {
        if ((int64_t)row.size() != ddim)
            fail(HANN_ERR_FORMAT, strf("coarse centroid of %zu values in an index of dimension %lld", row.size(),
                                       (long long)ddim));
        S.coarse.insert(S.coarse.end(), row.begin(), row.end());
    }
    S.have_codebooks = have_cb;
    if (have_cb) {
        if ((int64_t)CB.size() != dm)
            fail(HANN_ERR_FORMAT, strf("%zu codebooks for %lld subquantizers", CB.size(), (long long)dm));
        for (const auto &cb : CB) S.pmax = std::max<int64_t>(S.pmax, (int64_t)cb.size());
        S.cb_len.resize((size_t)dm);
        S.codebooks.assign((size_t)(dm * S.pmax * s), 0.0f);
        for (int64_t i = 0; i < dm; ++i) {
            S.cb_len[(size_t)i] = (int64_t)CB[(size_t)i].size();
            for (size_t j = 0; j < CB[(size_t)i].size(); ++j) {
                const auto &row = CB[(size_t)i][j];
                if ((int64_t)row.size() != s)
                    fail(HANN_ERR_FORMAT, strf("codeword of %zu values for subvectors of %lld", row.size(), (long long)s));
                std::copy(row.begin(), row.end(), S.codebooks.begin() + (size_t)((i * S.pmax + (int64_t)j) * s));
            }
        }
    } else {
        S.cb_len.assign((size_t)dm, 0);
    }
    S.lists.resize((size_t)S.kc);
    for (auto &pr : lists_in) {
        const int64_t c = pr.first;
        if (c < 0 || c >= S.kc) fail(HANN_ERR_FORMAT, strf("serialized list %lld has no coarse centroid", (long long)c));
        ListData &L = S.lists[(size_t)c];
        L = ListData{};   // a repeated key replaces the earlier list, as in a Go map
        L.exists = true;
        const size_t n = pr.second.size();
        L.ids.resize(n);
        L.has_codes.resize(n);
        L.codes.assign(n * (size_t)dm, 0);
        L.vecs.resize(n * (size_t)ddim);
        for (size_t r = 0; r < n; ++r) {
            const GobEntry &E = pr.second[r];
            if ((int64_t)E.vec.size() != ddim)
                fail(HANN_ERR_FORMAT, strf("entry %lld has %zu values, expected %lld", (long long)E.id, E.vec.size(),
                                           (long long)ddim));
            L.ids[r] = E.id;
            std::copy(E.vec.begin(), E.vec.end(), L.vecs.begin() + r * (size_t)ddim);
            // Search decodes an entry's codes only when it holds m of them
            // and every one lies in its codebook; otherwise it ranks the
            // entry by its vector (pqivf/index.go:906-928, decodePQCode
            // 678-690). Such entries are kept without codes. A negative code
            // passes decodePQCode's bound check and then indexes out of
            // range, so a file holding one is rejected.
            bool coded = (int64_t)E.codes.size() == dm && have_cb;
            for (size_t j = 0; j < E.codes.size(); ++j) {
                if (E.codes[j] < 0)
                    fail(HANN_ERR_FORMAT, strf("entry %lld has negative code %lld", (long long)E.id,
                                               (long long)E.codes[j]));
                if (j >= (size_t)dm || E.codes[j] >= S.cb_len[j]) coded = false;
            }
            L.has_codes[r] = coded ? 1 : 0;
            if (coded)
                for (size_t j = 0; j < E.codes.size(); ++j) L.codes[r * (size_t)dm + j] = (int32_t)E.codes[j];
        }
    }
    // idToCluster rebuilt from the lists (pqivf/index.go:1076-1081)
    for (size_t c = 0; c < S.lists.size(); ++c)
        for (int64_t id : S.lists[c].ids) S.id_to_cluster[id] = (int64_t)c;
    S.pending.dim = ddim;
    for (auto &pr : pend_in) {
        if ((int64_t)pr.second.size() != ddim)
            fail(HANN_ERR_FORMAT, strf("pending vector %lld has %zu values, expected %lld", (long long)pr.first,
                                       pr.second.size(), (long long)ddim));
        if (!S.pending.count(pr.first)) S.pending.put(pr.first, pr.second.data());
    }
    install(std::move(S));
}

// Load (pqivf/index.go:1094-1100)
void PQIVF::gob_load(const void *buf, size_t len) {
    const unsigned char *b = reinterpret_cast<const unsigned char *>(buf);
    Stream st{In{b, b + len}, {}};
    In m{nullptr, nullptr};
    const int id = st.next(m);
    if (id < 64 || st.type(id).kind != K_GOBENC) bad("the value is not an encoded index");
    if (m.u() != 0) bad("invalid singleton value");
    const std::string inner = m.str();
    gob_decode(inner.data(), inner.size());
}

}  // namespace hann
