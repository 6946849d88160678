// bsdata_io.h — a small JSON value type and the .bsdata reader.
//
// Format contract: benchmark/bsdata.py:1-17, 58-270 and
// simulator/csim/src/bsdata_reader.h:179-335 (JSON header, NUL, binary;
// "_bin", "_array_bin", "_broadcast_pack", inline ints / hex strings / lists,
// "cases" or legacy "layers", "bitlength" or legacy "n_vectors").
// nlohmann/json, which the reference uses, is not available offline, so the
// JSON parser and writer here are this library's own.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace bsz {

struct Json {
    enum Kind { Null, Bool, Int, Float, Str, Arr, Obj } k = Null;
    bool b = false;
    int64_t i = 0;
    bool is_unsigned_big = false;  // integer above INT64_MAX (stored in u)
    uint64_t u = 0;
    double f = 0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;  // insertion order kept

    const Json* get(const std::string& key) const;
    bool has(const std::string& key) const { return get(key) != nullptr; }
    int64_t as_int(int64_t dflt = 0) const;
};

Json json_parse(const std::string& text);       // throws std::runtime_error
std::string json_escape(const std::string& s);

struct CaseData {
    std::string name;
    int64_t bitlength = 0;
    std::unordered_map<std::string, int64_t> params;
    std::unordered_map<std::string, std::vector<uint64_t>> inputs;
    std::map<std::string, std::map<int64_t, std::vector<uint64_t>>> input_arrays;
    std::map<std::string, std::vector<bool>> broadcast;
    int iterations = 1;
    std::map<std::string, std::string> feedback;
    // 0: clean values (no bit at or above the bitlength, DSL.md:273-275);
    // 64 or 32: the padding bits of bsim_cuda's padded streams with words of
    // that width, which ~, ONES, + and << leave set and >> brings back.
    int bsim_padding = 0;
};

// All cases of a .bsdata file (read_bsdata_multi, bsdata_reader.h:262-299).
std::vector<CaseData> read_bsdata_file(const std::string& path);
// A JSON case from stdin (parse_json_input, bsdata_reader.h:303-335).
CaseData case_from_json(const Json& j);
// Words of a JSON value: hex string (optional "0x", "-" two's complement) or
// integer (word_utils.h:14-64, 118-133).
std::vector<uint64_t> json_value_words(const Json& v, size_t n_words);
// "0x..." without leading zeros, "0x0" for zero (word_utils.h:66-89).
std::string words_to_hex(const uint64_t* w, size_t n);
// --dump-input view of a case (bsdata_reader.h:338-359).
std::string dump_case_json(const CaseData& c, int indent);

}  // namespace bsz
