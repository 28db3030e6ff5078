#include "dictionary.hpp"

#include <cctype>
#include <cstdlib>

namespace mrfg {

double Value::asNumber() const
{
    if (kind != Number) throw DictError("expected a number, found '" + text + "'");
    return number;
}

const std::string &Value::asWord() const
{
    if (kind != Word && kind != String) throw DictError("expected a word");
    return text;
}

namespace {

struct Tok {
    enum Kind { End, Punct, Word, String } kind = End;
    char punct = 0;
    std::string text;
    int line = 0;
};

class Lexer {
public:
    explicit Lexer(const std::string &s) : s_(s) {}

    Tok next()
    {
        if (have_) {
            have_ = false;
            return peeked_;
        }
        return lex();
    }
    const Tok &peek()
    {
        if (!have_) {
            peeked_ = lex();
            have_ = true;
        }
        return peeked_;
    }
    void skipLine()
    {
        have_ = false;
        while (i_ < s_.size() && s_[i_] != '\n') ++i_;
    }
    int line() const { return line_; }

private:
    Tok lex()
    {
        skipSpace();
        Tok t;
        t.line = line_;
        if (i_ >= s_.size()) return t;
        char c = s_[i_];
        if (c == '{' || c == '}' || c == '(' || c == ')' || c == ';' || c == '[' || c == ']') {
            ++i_;
            t.kind = Tok::Punct;
            t.punct = c == '[' ? '(' : c == ']' ? ')' : c;
            return t;
        }
        if (c == '"') {
            ++i_;
            t.kind = Tok::String;
            while (i_ < s_.size() && s_[i_] != '"') {
                if (s_[i_] == '\\' && i_ + 1 < s_.size()) ++i_;
                if (s_[i_] == '\n') ++line_;
                t.text += s_[i_++];
            }
            if (i_ >= s_.size()) throw DictError("unterminated string at line " + std::to_string(t.line));
            ++i_;
            return t;
        }
        t.kind = Tok::Word;
        while (i_ < s_.size()) {
            char d = s_[i_];
            if (std::isspace((unsigned char)d) || d == '{' || d == '}' || d == '(' || d == ')' ||
                d == ';' || d == '"' || d == '[' || d == ']')
                break;
            if (d == '/' && i_ + 1 < s_.size() && (s_[i_ + 1] == '/' || s_[i_ + 1] == '*')) break;
            t.text += d;
            ++i_;
        }
        return t;
    }

    void skipSpace()
    {
        for (;;) {
            while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) {
                if (s_[i_] == '\n') ++line_;
                ++i_;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
                continue;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                i_ += 2;
                while (i_ + 1 < s_.size() && !(s_[i_] == '*' && s_[i_ + 1] == '/')) {
                    if (s_[i_] == '\n') ++line_;
                    ++i_;
                }
                if (i_ + 1 >= s_.size()) throw DictError("unterminated comment");
                i_ += 2;
                continue;
            }
            break;
        }
    }

    const std::string &s_;
    size_t i_ = 0;
    int line_ = 1;
    bool have_ = false;
    Tok peeked_;
};

Value wordValue(const Tok &t)
{
    Value v;
    if (t.kind == Tok::String) {
        v.kind = Value::String;
        v.text = t.text;
        return v;
    }
    v.text = t.text;
    char *end = nullptr;
    double d = std::strtod(t.text.c_str(), &end);
    if (!t.text.empty() && end && *end == '\0') {
        v.kind = Value::Number;
        v.number = d;
    } else {
        v.kind = Value::Word;
    }
    return v;
}

std::shared_ptr<Dict> parseEntries(Lexer &lx, bool braced, const std::string &name);

Value parseList(Lexer &lx)
{
    // '(' already consumed
    Value v;
    v.kind = Value::List;
    for (;;) {
        Tok t = lx.next();
        if (t.kind == Tok::End) throw DictError("unterminated list");
        if (t.kind == Tok::Punct) {
            if (t.punct == ')') return v;
            if (t.punct == '(') {
                v.items.push_back(parseList(lx));
                continue;
            }
            if (t.punct == '{') {
                Value d;
                d.kind = Value::Dictionary;
                std::string nm = (!v.items.empty() && v.items.back().isWord()) ? v.items.back().text : "";
                d.dict = parseEntries(lx, true, nm);
                v.items.push_back(d);
                continue;
            }
            if (t.punct == ';') continue;
            throw DictError("unexpected '" + std::string(1, t.punct) + "' in list at line " +
                            std::to_string(t.line));
        }
        v.items.push_back(wordValue(t));
    }
}

std::shared_ptr<Dict> parseEntries(Lexer &lx, bool braced, const std::string &name)
{
    auto d = std::make_shared<Dict>();
    d->name = name;
    for (;;) {
        Tok t = lx.next();
        if (t.kind == Tok::End) {
            if (braced) throw DictError("unterminated dictionary '" + name + "'");
            return d;
        }
        if (t.kind == Tok::Punct) {
            if (t.punct == '}' && braced) return d;
            if (t.punct == ';') continue;   // stray ';' after a closing brace
            throw DictError("unexpected '" + std::string(1, t.punct) + "' at line " +
                            std::to_string(t.line));
        }
        if (t.kind == Tok::Word && !t.text.empty() && t.text[0] == '#') {
            lx.skipLine();                  // #include / #inputMode directives
            continue;
        }
        Entry e;
        e.key = t.text;
        const Tok &p = lx.peek();
        if (p.kind == Tok::Punct && p.punct == '{') {
            lx.next();
            e.dict = parseEntries(lx, true, e.key);
            d->entries.push_back(std::move(e));
            continue;
        }
        for (;;) {
            Tok v = lx.next();
            if (v.kind == Tok::End) throw DictError("missing ';' after '" + e.key + "'");
            if (v.kind == Tok::Punct) {
                if (v.punct == ';') break;
                if (v.punct == '(') {
                    e.tokens.push_back(parseList(lx));
                    continue;
                }
                if (v.punct == '{') {
                    Value dv;
                    dv.kind = Value::Dictionary;
                    dv.dict = parseEntries(lx, true, e.key);
                    e.tokens.push_back(dv);
                    continue;
                }
                throw DictError("unexpected '" + std::string(1, v.punct) + "' in entry '" + e.key +
                                "' at line " + std::to_string(v.line));
            }
            e.tokens.push_back(wordValue(v));
        }
        d->entries.push_back(std::move(e));
    }
}

}  // namespace

const Entry *Dict::find(const std::string &key) const
{
    for (auto it = entries.rbegin(); it != entries.rend(); ++it)
        if (it->key == key) return &*it;
    return nullptr;
}

bool Dict::isDict(const std::string &key) const
{
    const Entry *e = find(key);
    return e && e->dict;
}

const Entry &Dict::lookup(const std::string &key) const
{
    const Entry *e = find(key);
    if (!e) throw DictError("keyword '" + key + "' is undefined in dictionary '" + name + "'");
    return *e;
}

const Dict &Dict::subDict(const std::string &key) const
{
    const Entry &e = lookup(key);
    if (!e.dict) throw DictError("keyword '" + key + "' is not a sub-dictionary of '" + name + "'");
    return *e.dict;
}

double Dict::number(const std::string &key) const
{
    const Entry &e = lookup(key);
    if (e.tokens.empty()) throw DictError("keyword '" + key + "' has no value");
    return e.tokens.front().asNumber();
}

double Dict::numberOrDefault(const std::string &key, double def) const
{
    return found(key) ? number(key) : def;
}

std::string Dict::word(const std::string &key) const
{
    const Entry &e = lookup(key);
    if (e.tokens.empty()) throw DictError("keyword '" + key + "' has no value");
    return e.tokens.front().text;
}

std::string Dict::wordOrDefault(const std::string &key, const std::string &def) const
{
    return found(key) ? word(key) : def;
}

const std::vector<Value> &Dict::list(const std::string &key) const
{
    const Entry &e = lookup(key);
    if (e.tokens.empty() || !e.tokens.front().isList())
        throw DictError("keyword '" + key + "' is not a list");
    return e.tokens.front().items;
}

Dict parseDict(const std::string &text)
{
    Lexer lx(text);
    return *parseEntries(lx, false, "");
}

}  // namespace mrfg
