// bsdata_io.cpp — JSON parsing and the .bsdata reader (see bsdata_io.h).
#include "bsdata_io.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace bsz {

const Json* Json::get(const std::string& key) const {
    if (k != Obj) return nullptr;
    for (auto& kv : o)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

int64_t Json::as_int(int64_t dflt) const {
    if (k == Int) return i;
    if (k == Float) return (int64_t)f;
    if (k == Bool) return b ? 1 : 0;
    return dflt;
}

namespace {

struct JP {
    const char* p;
    const char* e;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("JSON parse error: ") + what);
    }
    void ws() { while (p < e && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p; }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if ((size_t)(e - p) >= n && std::memcmp(p, w, n) == 0) { p += n; return true; }
        return false;
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string str() {
        if (p >= e || *p != '"') fail("expected string");
        ++p;
        std::string out;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) fail("bad escape");
                char c = *p++;
                switch (c) {
                    case 'n': out += '\n'; break; case 't': out += '\t'; break;
                    case 'r': out += '\r'; break; case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (e - p < 4) fail("bad \\u");
                        uint32_t cp = (uint32_t)std::stoul(std::string(p, 4), nullptr, 16);
                        p += 4;
                        if (cp >= 0xD800 && cp < 0xDC00 && e - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                            uint32_t lo = (uint32_t)std::stoul(std::string(p + 2, 4), nullptr, 16);
                            p += 6;
                            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        }
                        utf8(out, cp);
                        break;
                    }
                    default: out += c;
                }
            } else {
                out += *p++;
            }
        }
        if (p >= e) fail("unterminated string");
        ++p;
        return out;
    }
    Json val() {
        ws();
        if (p >= e) fail("unexpected end");
        Json j;
        char c = *p;
        if (c == '{') {
            ++p; j.k = Json::Obj; ws();
            if (p < e && *p == '}') { ++p; return j; }
            for (;;) {
                ws();
                std::string key = str();
                ws();
                if (p >= e || *p != ':') fail("expected ':'");
                ++p;
                j.o.emplace_back(std::move(key), val());
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            ++p; j.k = Json::Arr; ws();
            if (p < e && *p == ']') { ++p; return j; }
            for (;;) {
                j.a.push_back(val());
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            j.k = Json::Str; j.s = str();
        } else if (lit("true")) { j.k = Json::Bool; j.b = true; }
        else if (lit("false")) { j.k = Json::Bool; j.b = false; }
        else if (lit("null")) { j.k = Json::Null; }
        else {
            const char* b = p;
            if (*p == '-') ++p;
            bool fl = false;
            while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) {
                if (*p == '.' || *p == 'e' || *p == 'E') fl = true;
                ++p;
            }
            std::string num(b, p);
            if (num.empty() || num == "-") fail("bad value");
            if (fl) { j.k = Json::Float; j.f = std::stod(num); }
            else {
                j.k = Json::Int;
                errno = 0;
                if (num[0] == '-') j.i = std::stoll(num);
                else {
                    unsigned long long u = std::stoull(num);
                    if (u > (unsigned long long)INT64_MAX) { j.is_unsigned_big = true; j.u = u; j.i = (int64_t)u; }
                    else j.i = (int64_t)u;
                }
            }
        }
        return j;
    }
};

std::vector<uint64_t> bytes_to_words(const uint8_t* data, size_t nbytes, size_t n_words) {
    std::vector<uint64_t> w(n_words, 0);
    size_t cp = std::min(nbytes, n_words * 8);
    if (cp) std::memcpy(w.data(), data, cp);
    return w;
}

size_t words_for(int64_t bitlength) {
    return bitlength <= 0 ? 1 : (size_t)((bitlength + 63) / 64);
}

struct Blob {
    const uint8_t* p;
    size_t n;
};

[[noreturn]] void bad_section(const std::string& key, const char* what) {
    throw std::runtime_error("Invalid bsdata section '" + key + "': " + what);
}

// Integer field i of a section descriptor (an array of non-negative ints).
uint64_t desc_field(const Json& r, size_t i, const std::string& key) {
    if (r.k != Json::Arr || r.a.size() <= i) bad_section(key, "descriptor array too short");
    const Json& x = r.a[i];
    if (x.k != Json::Int || x.is_unsigned_big) bad_section(key, "descriptor field is not an integer");
    const int64_t v = x.as_int(-1);
    if (v < 0) bad_section(key, "negative offset, size or count");
    return (uint64_t)v;
}

// [off, off + len) lies inside the binary blob (no overflow on huge offsets).
bool in_blob(uint64_t off, uint64_t len, const Blob& bin) { return off <= bin.n && len <= bin.n - off; }

std::vector<uint64_t> decode_stream_value(const Json& v, const Blob& bin, size_t nw, const std::string& key) {
    if (v.k == Json::Obj) {
        const Json* r = v.get("_bin");
        if (!r) return std::vector<uint64_t>(nw, 0);
        const uint64_t off = desc_field(*r, 0, key), nb = desc_field(*r, 1, key);
        if (!in_blob(off, nb, bin)) bad_section(key, "_bin range lies outside the binary part");
        return bytes_to_words(bin.p + off, nb, nw);
    }
    return json_value_words(v, nw);
}

void decode_section(const Json& sec, const Blob& bin, size_t nw, CaseData& c, bool scalars) {
    if (sec.k != Json::Obj) return;
    for (auto& kv : sec.o) {
        const std::string& k = kv.first;
        const Json& v = kv.second;
        if (v.k == Json::Obj && v.has("_broadcast_pack")) {
            const Json& r = *v.get("_broadcast_pack");
            const uint64_t count = desc_field(r, 0, k), off = desc_field(r, 1, k);
            if (!in_blob(off, (count + 7) / 8, bin)) bad_section(k, "_broadcast_pack bits lie outside the binary part");
            std::vector<bool> bv((size_t)count);
            for (uint64_t i = 0; i < count; ++i) bv[i] = (bin.p[off + i / 8] >> (i % 8)) & 1;
            c.broadcast[k] = std::move(bv);
        } else if (v.k == Json::Obj && v.has("_array_bin")) {
            const Json& r = *v.get("_array_bin");
            const uint64_t count = desc_field(r, 0, k), off = desc_field(r, 1, k), stride = desc_field(r, 2, k);
            if (count > (1ull << 31)) bad_section(k, "_array_bin count too large");
            if (count && (stride > bin.n || !in_blob(off, count * stride, bin)))
                bad_section(k, "_array_bin elements lie outside the binary part");
            auto& arr = c.input_arrays[k];
            for (uint64_t i = 0; i < count; ++i) arr[(int64_t)i] = bytes_to_words(bin.p + off + i * stride, stride, nw);
        } else if (v.k == Json::Obj && v.has("_bin")) {
            if (scalars) c.inputs[k] = decode_stream_value(v, bin, nw, k);
        } else if (v.k == Json::Arr) {
            auto& arr = c.input_arrays[k];
            for (size_t i = 0; i < v.a.size(); ++i) arr[(int64_t)i] = json_value_words(v.a[i], nw);
        } else if (v.k == Json::Obj) {
            auto& arr = c.input_arrays[k];
            for (auto& e : v.o) arr[std::stoll(e.first)] = decode_stream_value(e.second, bin, nw, k);
        } else if (scalars) {
            c.inputs[k] = json_value_words(v, nw);
        }
    }
}

CaseData parse_case(const Json& layer, const Blob& bin) {
    CaseData c;
    const Json* bl = layer.get("bitlength");
    if (!bl) bl = layer.get("n_vectors");
    c.bitlength = bl ? bl->as_int() : 0;
    size_t nw = words_for(c.bitlength);
    if (const Json* p = layer.get("params"))
        for (auto& kv : p->o) c.params[kv.first] = kv.second.as_int();
    if (const Json* in = layer.get("inputs")) decode_section(*in, bin, nw, c, true);
    if (const Json* ia = layer.get("input_arrays")) decode_section(*ia, bin, nw, c, false);
    if (const Json* it = layer.get("iterations")) c.iterations = (int)it->as_int(1);
    if (const Json* fb = layer.get("feedback"))
        for (auto& kv : fb->o) c.feedback[kv.first] = kv.second.s;
    return c;
}

}  // namespace

Json json_parse(const std::string& text) {
    JP jp{text.data(), text.data() + text.size()};
    Json j = jp.val();
    jp.ws();
    if (jp.p != jp.e) jp.fail("trailing characters");
    return j;
}

std::string json_escape(const std::string& s) {
    std::string o;
    o.reserve(s.size() + 2);
    o += '"';
    for (char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\t': o += "\\t"; break;
            case '\r': o += "\\r"; break;
            default:
                if ((unsigned char)c < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", c); o += b; }
                else o += c;
        }
    }
    o += '"';
    return o;
}

std::vector<uint64_t> json_value_words(const Json& v, size_t n_words) {
    std::vector<uint64_t> w(n_words, 0);
    if (n_words == 0) return w;
    if (v.k == Json::Str) {
        const std::string& s = v.s;
        size_t st = 0;
        bool neg = false;
        if (s.size() >= 2 && s[0] == '0' && (s[1] == 'x' || s[1] == 'X')) st = 2;
        if (s.size() > st && s[st] == '-') { neg = true; ++st; }
        // A leading "-0x" (Python's hex of a negative) is read the same way.
        if (s.size() >= st + 2 && s[st] == '0' && (s[st + 1] == 'x' || s[st + 1] == 'X')) st += 2;
        size_t end = s.size();
        size_t wi = 0;
        while (end > st && wi < n_words) {
            size_t b = end >= st + 16 ? end - 16 : st;
            uint64_t word = 0;
            for (size_t i = b; i < end; ++i) {
                char c = s[i];
                uint64_t nib = (c >= '0' && c <= '9') ? c - '0' : (c >= 'a' && c <= 'f') ? c - 'a' + 10
                             : (c >= 'A' && c <= 'F') ? c - 'A' + 10 : 0;
                word = (word << 4) | nib;
            }
            w[wi++] = word;
            end = b;
        }
        if (neg) {
            uint64_t carry = 1;
            for (size_t i = 0; i < n_words; ++i) {
                uint64_t x = ~w[i], y = x + carry;
                carry = y < x;
                w[i] = y;
            }
        }
        return w;
    }
    if (v.k == Json::Int) {
        if (v.is_unsigned_big) { w[0] = v.u; return w; }
        w[0] = (uint64_t)v.i;
        if (v.i < 0) for (size_t i = 1; i < n_words; ++i) w[i] = ~0ull;
        return w;
    }
    if (v.k == Json::Bool) { w[0] = v.b ? 1 : 0; return w; }
    return w;
}

std::string words_to_hex(const uint64_t* w, size_t n) {
    size_t top = n;
    while (top > 0 && w[top - 1] == 0) --top;
    if (top == 0) return "0x0";
    std::string out;
    out.reserve(2 + top * 16);
    out += "0x";
    char buf[20];
    std::snprintf(buf, sizeof buf, "%llx", (unsigned long long)w[top - 1]);
    out += buf;
    static const char* hexd = "0123456789abcdef";
    for (size_t i = top - 1; i > 0; --i) {
        uint64_t x = w[i - 1];
        char t[16];
        for (int d = 15; d >= 0; --d) { t[d] = hexd[x & 15]; x >>= 4; }
        out.append(t, 16);
    }
    return out;
}

namespace {

// A read-only memory map of a whole file (the reference reads the file into
// a vector and copies each stream out again, bsdata_reader.h:262-299; here
// streams are decoded straight from the page cache).
struct MappedFile {
    const uint8_t* p = nullptr;
    size_t n = 0;
    int fd = -1;
    explicit MappedFile(const std::string& path) {
        fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) throw std::runtime_error("Cannot open bsdata file: " + path);
        struct stat st;
        if (::fstat(fd, &st) != 0) { ::close(fd); throw std::runtime_error("Cannot stat bsdata file: " + path); }
        n = (size_t)st.st_size;
        if (n) {
            void* m = ::mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) { ::close(fd); throw std::runtime_error("Cannot map bsdata file: " + path); }
            (void)::madvise(m, n, MADV_SEQUENTIAL);
            p = static_cast<const uint8_t*>(m);
        }
    }
    ~MappedFile() {
        if (p) ::munmap(const_cast<uint8_t*>(p), n);
        if (fd >= 0) ::close(fd);
    }
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
};

}  // namespace

std::vector<CaseData> read_bsdata_file(const std::string& path) {
    MappedFile f(path);
    const uint8_t* sep = f.n ? static_cast<const uint8_t*>(std::memchr(f.p, 0, f.n)) : nullptr;
    if (!sep) throw std::runtime_error("Invalid bsdata file (no NUL separator): " + path);
    const size_t sp = (size_t)(sep - f.p);
    Json header = json_parse(std::string(reinterpret_cast<const char*>(f.p), sp));
    Blob bin{f.p + sp + 1, f.n - sp - 1};
    std::vector<CaseData> out;
    const Json* multi = header.get("cases");
    if (!multi) multi = header.get("layers");
    if (multi) {
        for (size_t i = 0; i < multi->a.size(); ++i) out.push_back(parse_case(multi->a[i], bin));
        if (out.size() > 1)
            for (size_t i = 0; i < out.size(); ++i) out[i].name = path + "[" + std::to_string(i) + "]";
        else if (!out.empty()) out[0].name = path;
    } else {
        out.push_back(parse_case(header, bin));
        out[0].name = path;
    }
    return out;
}

CaseData case_from_json(const Json& j) {
    CaseData c;
    const Json* bl = j.get("bitlength");
    if (!bl) bl = j.get("n_vectors");
    c.bitlength = bl ? bl->as_int() : 0;
    size_t nw = words_for(c.bitlength);
    if (const Json* in = j.get("inputs"))
        for (auto& kv : in->o) c.inputs[kv.first] = json_value_words(kv.second, nw);
    if (const Json* p = j.get("params"))
        for (auto& kv : p->o) c.params[kv.first] = kv.second.as_int();
    if (const Json* ia = j.get("input_arrays")) {
        for (auto& kv : ia->o) {
            auto& arr = c.input_arrays[kv.first];
            if (kv.second.k == Json::Arr)
                for (size_t i = 0; i < kv.second.a.size(); ++i) arr[(int64_t)i] = json_value_words(kv.second.a[i], nw);
            else if (kv.second.k == Json::Obj)
                for (auto& e : kv.second.o) arr[std::stoll(e.first)] = json_value_words(e.second, nw);
        }
    }
    if (const Json* it = j.get("iterations")) c.iterations = (int)it->as_int(1);
    if (const Json* fb = j.get("feedback"))
        for (auto& kv : fb->o) c.feedback[kv.first] = kv.second.s;
    return c;
}

std::string dump_case_json(const CaseData& c, int indent) {
    std::string pad(indent, ' '), pad2(indent * 2, ' '), pad3(indent * 3, ' ');
    std::string nl = indent ? "\n" : "";
    std::string o = "{" + nl;
    o += pad + "\"bitlength\": " + std::to_string(c.bitlength) + "," + nl;
    o += pad + "\"inputs\": {";
    bool first = true;
    std::map<std::string, const std::vector<uint64_t>*> sorted_in;
    for (auto& kv : c.inputs) sorted_in[kv.first] = &kv.second;
    for (auto& kv : sorted_in) {
        o += (first ? nl : "," + nl) + pad2 + json_escape(kv.first) + ": \"" + words_to_hex(kv.second->data(), kv.second->size()) + "\"";
        first = false;
    }
    o += (first ? "" : nl + pad) + "}," + nl;
    o += pad + "\"params\": {";
    first = true;
    std::map<std::string, int64_t> sorted_p(c.params.begin(), c.params.end());
    for (auto& kv : sorted_p) {
        o += (first ? nl : "," + nl) + pad2 + json_escape(kv.first) + ": " + std::to_string(kv.second);
        first = false;
    }
    o += (first ? "" : nl + pad) + "}," + nl;
    o += pad + "\"input_arrays\": {";
    first = true;
    for (auto& kv : c.input_arrays) {
        o += (first ? nl : "," + nl) + pad2 + json_escape(kv.first) + ": {";
        bool f2 = true;
        for (auto& e : kv.second) {
            o += (f2 ? nl : "," + nl) + pad3 + "\"" + std::to_string(e.first) + "\": \"" +
                 words_to_hex(e.second.data(), e.second.size()) + "\"";
            f2 = false;
        }
        o += (f2 ? "" : nl + pad2) + "}";
        first = false;
    }
    o += (first ? "" : nl + pad) + "}" + nl + "}";
    return o;
}

}  // namespace bsz
