// OpenFOAM dictionary subset: the syntax of fvSolution / fvSchemes /
// multiRegionProperties / regionInterfaceProperties files.
//
//   key value ... ;          entry with a token stream
//   key { ... }              sub-dictionary
//   ( a b { ... } ( ... ) )  lists, possibly holding dictionaries and lists
//   // and /* */ comments, FoamFile header (ignored like any other entry)
//
// Values are kept as a small tree; lookups follow OpenFOAM's semantics
// (last definition of a key wins, lookupOrDefault for optional keys).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mrfg {

struct DictError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Dict;

struct Value {
    enum Kind { Word, Number, String, List, Dictionary } kind = Word;
    std::string text;                   // Word / String / Number text
    double number = 0.0;                // Number
    std::vector<Value> items;           // List
    std::shared_ptr<Dict> dict;         // Dictionary

    bool isWord() const { return kind == Word; }
    bool isNumber() const { return kind == Number; }
    bool isList() const { return kind == List; }
    bool isDict() const { return kind == Dictionary; }
    double asNumber() const;
    const std::string &asWord() const;
};

struct Entry {
    std::string key;
    std::vector<Value> tokens;          // stream entry (empty for a dictionary)
    std::shared_ptr<Dict> dict;         // sub-dictionary entry
};

struct Dict {
    std::string name;
    std::vector<Entry> entries;

    const Entry *find(const std::string &key) const;    // last definition wins
    bool found(const std::string &key) const { return find(key) != nullptr; }
    bool isDict(const std::string &key) const;
    const Dict &subDict(const std::string &key) const;
    const Entry &lookup(const std::string &key) const;
    double number(const std::string &key) const;
    double numberOrDefault(const std::string &key, double def) const;
    std::string word(const std::string &key) const;
    std::string wordOrDefault(const std::string &key, const std::string &def) const;
    // a list entry "key ( ... );"
    const std::vector<Value> &list(const std::string &key) const;
};

// Parse a whole file's text into its top-level dictionary.
Dict parseDict(const std::string &text);

}  // namespace mrfg
