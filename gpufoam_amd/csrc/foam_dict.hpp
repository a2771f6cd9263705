// Minimal reader for the OpenFOAM dictionary subset used by thermo,
// reactions and chemistryProperties files: `key value...;`, `key { ... }`,
// lists `( ... )` with an optional size prefix, "strings", // and /* */
// comments.  Words may contain balanced parentheses (specie names such as
// CH2(S)), following OpenFOAM's word tokenizer.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace gfc {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Token {
    enum Kind { Word, String, Punct } kind;
    std::string text; // for Punct: one of ( ) ;
};

struct Dict;

struct Entry {
    std::string key;
    std::vector<Token> tokens;   // value tokens (stream entry)
    std::unique_ptr<Dict> dict;  // sub-dictionary entry
};

struct Dict {
    std::vector<Entry> entries; // file order preserved

    const Entry* find(const std::string& key) const;
    const Dict* subDict(const std::string& key) const; // nullptr if absent
    const Dict& requireDict(const std::string& key) const;
    const std::vector<Token>& requireTokens(const std::string& key) const;
    bool found(const std::string& key) const { return find(key) != nullptr; }

    double scalar(const std::string& key) const;
    double scalarOr(const std::string& key, double dflt) const;
    std::string word(const std::string& key) const;
    // `key N? ( w1 w2 ... );`
    std::vector<std::string> wordList(const std::string& key) const;
    // `key N? ( x1 x2 ... );`
    std::vector<double> scalarList(const std::string& key) const;
    // `key N? ( (w x) (w x) ... );`
    std::vector<std::pair<std::string, double>>
    wordScalarList(const std::string& key) const;
};

Dict parseDict(const std::string& text);
double toScalar(const std::string& s);

} // namespace gfc
