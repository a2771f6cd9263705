#include "foam_dict.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>

namespace gfc {

namespace {

class Lexer {
public:
    explicit Lexer(const std::string& s) : s_(s) {}

    bool next(Token& t) {
        skipSpaceAndComments();
        if (i_ >= s_.size()) return false;
        const char ch = s_[i_];
        if (ch == '(' || ch == ')' || ch == ';' || ch == '{' || ch == '}') {
            t = Token{Token::Punct, std::string(1, ch)};
            ++i_;
            return true;
        }
        if (ch == '"') {
            ++i_;
            std::string out;
            while (i_ < s_.size() && s_[i_] != '"') {
                if (s_[i_] == '\\' && i_ + 1 < s_.size()) ++i_;
                out.push_back(s_[i_++]);
            }
            if (i_ >= s_.size()) throw ParseError("unterminated string");
            ++i_;
            t = Token{Token::String, out};
            return true;
        }
        // word: stops at space, ';', '{', '}', '"' or an unbalanced ')'
        std::string out;
        int depth = 0;
        while (i_ < s_.size()) {
            const char c = s_[i_];
            if (std::isspace(static_cast<unsigned char>(c)) || c == ';' ||
                c == '{' || c == '}' || c == '"')
                break;
            if (c == '(') {
                ++depth;
            } else if (c == ')') {
                if (depth == 0) break;
                --depth;
            }
            out.push_back(c);
            ++i_;
        }
        t = Token{Token::Word, out};
        return true;
    }

private:
    void skipSpaceAndComments() {
        for (;;) {
            while (i_ < s_.size() &&
                   std::isspace(static_cast<unsigned char>(s_[i_])))
                ++i_;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
                continue;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                const size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos)
                    throw ParseError("unterminated comment");
                i_ = e + 2;
                continue;
            }
            break;
        }
    }

    const std::string& s_;
    size_t i_ = 0;
};

void parseInto(Dict& d, std::vector<Token>& toks, size_t& pos, bool nested) {
    while (pos < toks.size()) {
        const Token& t = toks[pos];
        if (t.kind == Token::Punct && t.text == "}") {
            if (!nested) throw ParseError("unexpected '}'");
            ++pos;
            return;
        }
        if (t.kind == Token::Punct && t.text == ";") { // stray ';'
            ++pos;
            continue;
        }
        if (t.kind == Token::Punct)
            throw ParseError("expected keyword, found '" + t.text + "'");
        Entry e;
        e.key = t.text;
        ++pos;
        if (!e.key.empty() && e.key[0] == '#') {
            // directive (#include "file", #includeEtc ...): one argument,
            // no terminating ';'; kept as an entry, not expanded
            if (pos < toks.size() && toks[pos].kind != Token::Punct)
                e.tokens.push_back(toks[pos++]);
            d.entries.push_back(std::move(e));
            continue;
        }
        if (pos < toks.size() && toks[pos].kind == Token::Punct &&
            toks[pos].text == "{") {
            ++pos;
            e.dict = std::make_unique<Dict>();
            parseInto(*e.dict, toks, pos, true);
        } else {
            int depth = 0;
            while (pos < toks.size()) {
                const Token& v = toks[pos];
                if (v.kind == Token::Punct) {
                    if (v.text == "(") ++depth;
                    if (v.text == ")") --depth;
                    if (v.text == ";" && depth <= 0) {
                        ++pos;
                        break;
                    }
                    if (v.text == "{" || v.text == "}")
                        throw ParseError("unexpected brace in entry '" +
                                         e.key + "'");
                }
                e.tokens.push_back(v);
                ++pos;
            }
        }
        d.entries.push_back(std::move(e));
    }
    if (nested) throw ParseError("missing '}'");
}

// Strip an optional size prefix and the outer parentheses of a list value.
std::vector<Token> listBody(const std::vector<Token>& v, const std::string& k) {
    size_t b = 0;
    if (b < v.size() && v[b].kind == Token::Word) ++b; // N
    if (b >= v.size() || v[b].kind != Token::Punct || v[b].text != "(")
        throw ParseError("entry '" + k + "' is not a list");
    if (v.back().kind != Token::Punct || v.back().text != ")")
        throw ParseError("entry '" + k + "' list not closed");
    return std::vector<Token>(v.begin() + b + 1, v.end() - 1);
}

} // namespace

double toScalar(const std::string& s) {
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE)
        throw ParseError("not a number: '" + s + "'");
    return v;
}

Dict parseDict(const std::string& text) {
    Lexer lx(text);
    std::vector<Token> toks;
    Token t;
    while (lx.next(t)) toks.push_back(t);
    Dict d;
    size_t pos = 0;
    parseInto(d, toks, pos, false);
    return d;
}

const Entry* Dict::find(const std::string& key) const {
    // OpenFOAM keeps the last definition of a repeated keyword
    const Entry* hit = nullptr;
    for (const auto& e : entries)
        if (e.key == key) hit = &e;
    return hit;
}

const Dict* Dict::subDict(const std::string& key) const {
    const Entry* e = find(key);
    return (e && e->dict) ? e->dict.get() : nullptr;
}

const Dict& Dict::requireDict(const std::string& key) const {
    const Dict* d = subDict(key);
    if (!d) throw ParseError("missing sub-dictionary '" + key + "'");
    return *d;
}

const std::vector<Token>& Dict::requireTokens(const std::string& key) const {
    const Entry* e = find(key);
    if (!e || e->dict) throw ParseError("missing entry '" + key + "'");
    return e->tokens;
}

double Dict::scalar(const std::string& key) const {
    const auto& v = requireTokens(key);
    if (v.size() != 1 || v[0].kind != Token::Word)
        throw ParseError("entry '" + key + "' is not a scalar");
    return toScalar(v[0].text);
}

double Dict::scalarOr(const std::string& key, double dflt) const {
    return found(key) ? scalar(key) : dflt;
}

std::string Dict::word(const std::string& key) const {
    const auto& v = requireTokens(key);
    if (v.size() != 1 || v[0].kind == Token::Punct)
        throw ParseError("entry '" + key + "' is not a word");
    return v[0].text;
}

std::vector<std::string> Dict::wordList(const std::string& key) const {
    std::vector<std::string> out;
    for (const auto& t : listBody(requireTokens(key), key)) {
        if (t.kind == Token::Punct)
            throw ParseError("entry '" + key + "' is not a word list");
        out.push_back(t.text);
    }
    return out;
}

std::vector<double> Dict::scalarList(const std::string& key) const {
    std::vector<double> out;
    for (const auto& t : listBody(requireTokens(key), key)) {
        if (t.kind != Token::Word)
            throw ParseError("entry '" + key + "' is not a scalar list");
        out.push_back(toScalar(t.text));
    }
    return out;
}

std::vector<std::pair<std::string, double>>
Dict::wordScalarList(const std::string& key) const {
    const auto body = listBody(requireTokens(key), key);
    std::vector<std::pair<std::string, double>> out;
    size_t i = 0;
    while (i < body.size()) {
        if (i + 3 >= body.size())
            throw ParseError("entry '" + key + "' truncated");
        if (body[i].kind != Token::Punct || body[i].text != "(" ||
            body[i + 1].kind == Token::Punct ||
            body[i + 2].kind != Token::Word || body[i + 3].kind != Token::Punct ||
            body[i + 3].text != ")")
            throw ParseError("entry '" + key + "' is not a (word scalar) list");
        out.emplace_back(body[i + 1].text, toScalar(body[i + 2].text));
        i += 4;
    }
    return out;
}

} // namespace gfc
