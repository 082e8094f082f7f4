// dictionary.C -- see dictionary.H.
#include "dictionary.H"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace hs {

bool Node::isNumber() const {
    if (kind != Word || text.empty()) return false;
    char* end = nullptr;
    errno = 0;
    std::strtod(text.c_str(), &end);
    return end && *end == '\0' && errno == 0;
}

double Node::number() const {
    if (!isNumber()) throw FatalError("expected a number, found '" + text + "' (line " + std::to_string(line) + ")");
    return std::strtod(text.c_str(), nullptr);
}

namespace {

struct Tok {
    enum K { Word, String, Punct, End } k;
    std::string s;
    int line;
};

class Lexer {
public:
    Lexer(const std::string& t, const std::string& name) : t_(t), name_(name) {}
    Tok next() {
        skip();
        if (i_ >= t_.size()) return { Tok::End, "", line_ };
        const char c = t_[i_];
        if (c == '{' || c == '}' || c == '(' || c == ')' || c == ';' || c == '[' || c == ']') {
            ++i_;
            return { Tok::Punct, std::string(1, c), line_ };
        }
        if (c == '"') {
            const int l = line_;
            std::string s;
            ++i_;
            while (i_ < t_.size() && t_[i_] != '"') {
                if (t_[i_] == '\n') ++line_;
                s += t_[i_++];
            }
            if (i_ >= t_.size()) throw FatalError(name_ + ": unterminated string (line " + std::to_string(l) + ")");
            ++i_;
            return { Tok::String, s, l };
        }
        std::string s;
        while (i_ < t_.size()) {
            const char d = t_[i_];
            if (std::isspace((unsigned char)d) || d == '{' || d == '}' || d == '(' || d == ')' || d == ';' ||
                d == '"' || d == '[' || d == ']')
                break;
            if (d == '/' && i_ + 1 < t_.size() && (t_[i_ + 1] == '/' || t_[i_ + 1] == '*')) break;
            s += d;
            ++i_;
        }
        return { Tok::Word, s, line_ };
    }

private:
    void skip() {
        while (i_ < t_.size()) {
            const char c = t_[i_];
            if (c == '\n') { ++line_; ++i_; continue; }
            if (std::isspace((unsigned char)c)) { ++i_; continue; }
            if (c == '/' && i_ + 1 < t_.size() && t_[i_ + 1] == '/') {
                while (i_ < t_.size() && t_[i_] != '\n') ++i_;
                continue;
            }
            if (c == '/' && i_ + 1 < t_.size() && t_[i_ + 1] == '*') {
                i_ += 2;
                while (i_ + 1 < t_.size() && !(t_[i_] == '*' && t_[i_ + 1] == '/')) {
                    if (t_[i_] == '\n') ++line_;
                    ++i_;
                }
                i_ += 2;
                continue;
            }
            break;
        }
    }
    const std::string& t_;
    std::string name_;
    size_t i_ = 0;
    int line_ = 1;
};

class Parser {
public:
    Parser(const std::string& text, const std::string& name) : lx_(text, name), name_(name) { adv(); }

    NodePtr parseDictBody(bool top) {
        auto d = std::make_shared<Node>();
        d->kind = Node::Dict;
        d->line = cur_.line;
        for (;;) {
            if (cur_.k == Tok::End) {
                if (!top) fail("unexpected end of file inside a dictionary");
                return d;
            }
            if (cur_.k == Tok::Punct && cur_.s == "}") {
                if (top) fail("unbalanced '}'");
                return d;
            }
            if (cur_.k == Tok::Punct && cur_.s == ";") { adv(); continue; }
            if (cur_.k != Tok::Word && cur_.k != Tok::String) fail("expected a keyword, found '" + cur_.s + "'");
            const std::string key = cur_.s;
            adv();
            if (!key.empty() && key[0] == '#') {
                // directives (#include "file", #includeEtc, #inputMode merge, ...) are
                // recorded with their argument and take no ';'
                std::vector<NodePtr> value;
                if (cur_.k == Tok::Word || cur_.k == Tok::String) value.push_back(parseItem());
                d->entries.emplace_back(key, value);
                continue;
            }
            std::vector<NodePtr> value;
            if (cur_.k == Tok::Punct && cur_.s == "{") {
                adv();
                NodePtr sub = parseDictBody(false);
                expect("}");
                value.push_back(sub);
                d->entries.emplace_back(key, value);
                continue;
            }
            while (!(cur_.k == Tok::Punct && cur_.s == ";")) {
                if (cur_.k == Tok::End) fail("missing ';' after '" + key + "'");
                if (cur_.k == Tok::Punct && cur_.s == "}") fail("missing ';' after '" + key + "'");
                value.push_back(parseItem());
            }
            adv();
            d->entries.emplace_back(key, value);
        }
    }

private:
    NodePtr parseItem() {
        auto n = std::make_shared<Node>();
        n->line = cur_.line;
        if (cur_.k == Tok::Punct && cur_.s == "(") {
            adv();
            n->kind = Node::List;
            while (!(cur_.k == Tok::Punct && cur_.s == ")")) {
                if (cur_.k == Tok::End) fail("unterminated list");
                n->items.push_back(parseItem());
            }
            adv();
            return n;
        }
        if (cur_.k == Tok::Punct && cur_.s == "{") {
            adv();
            NodePtr d = parseDictBody(false);
            expect("}");
            return d;
        }
        if (cur_.k == Tok::Punct && cur_.s == "[") {  // dimension set: keep as a list of words
            adv();
            n->kind = Node::List;
            while (!(cur_.k == Tok::Punct && cur_.s == "]")) {
                if (cur_.k == Tok::End) fail("unterminated dimension set");
                auto w = std::make_shared<Node>();
                w->text = cur_.s;
                w->line = cur_.line;
                n->items.push_back(w);
                adv();
            }
            adv();
            return n;
        }
        if (cur_.k == Tok::Punct) fail("unexpected '" + cur_.s + "'");
        n->kind = cur_.k == Tok::String ? Node::String : Node::Word;
        n->text = cur_.s;
        adv();
        return n;
    }
    void expect(const char* p) {
        if (!(cur_.k == Tok::Punct && cur_.s == p)) fail(std::string("expected '") + p + "'");
        adv();
    }
    void adv() { cur_ = lx_.next(); }
    [[noreturn]] void fail(const std::string& m) {
        throw FatalError(name_ + ":" + std::to_string(cur_.line) + ": " + m);
    }
    Lexer lx_;
    std::string name_;
    Tok cur_{ Tok::End, "", 0 };
};

}  // namespace

dictionary parseDictionary(const std::string& text, const std::string& name) {
    Parser p(text, name);
    return dictionary(p.parseDictBody(true), name);
}

dictionary readDictionary(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw FatalError("cannot open file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parseDictionary(ss.str(), path);
}

std::vector<double> numbers(const NodePtr& list, const std::string& what) {
    if (!list || list->kind != Node::List) throw FatalError(what + ": expected a list of numbers");
    std::vector<double> out;
    for (const auto& it : list->items) out.push_back(it->number());
    return out;
}

// ---------------------------------------------------------------------------
void dictionary::fatal(const std::string& what, const std::string& key) const {
    throw FatalError("keyword " + key + " " + what + " in dictionary " + name_);
}

bool dictionary::found(const std::string& key) const {
    for (const auto& e : node_->entries)
        if (e.first == key) return true;
    return false;
}

const std::vector<NodePtr>& dictionary::lookup(const std::string& key) const {
    for (const auto& e : node_->entries)
        if (e.first == key) return e.second;
    fatal("is undefined", key);
}

bool dictionary::isDict(const std::string& key) const {
    for (const auto& e : node_->entries)
        if (e.first == key) return e.second.size() == 1 && e.second[0]->kind == Node::Dict;
    return false;
}

dictionary dictionary::subDict(const std::string& key) const {
    const auto& v = lookup(key);
    if (v.size() != 1 || v[0]->kind != Node::Dict) fatal("is not a sub-dictionary", key);
    return dictionary(v[0], name_ + "/" + key);
}

dictionary dictionary::subOrEmptyDict(const std::string& key) const {
    return isDict(key) ? subDict(key) : dictionary();
}

std::vector<std::string> dictionary::toc() const {
    std::vector<std::string> k;
    for (const auto& e : node_->entries) k.push_back(e.first);
    return k;
}

double dictionary::scalar(const std::string& key) const {
    const auto& v = lookup(key);
    if (v.size() != 1 || !v[0]->isNumber()) fatal("is not a number", key);
    return v[0]->number();
}

double dictionary::lookupOrDefault(const std::string& key, double def) const {
    return found(key) ? scalar(key) : def;
}

int dictionary::label(const std::string& key) const {
    const double x = scalar(key);
    if (x != (double)(int)x) fatal("is not an integer", key);
    return (int)x;
}

int dictionary::lookupOrDefault(const std::string& key, int def) const { return found(key) ? label(key) : def; }

std::string dictionary::word(const std::string& key) const {
    const auto& v = lookup(key);
    if (v.size() != 1 || (v[0]->kind != Node::Word && v[0]->kind != Node::String)) fatal("is not a word", key);
    return v[0]->text;
}

std::string dictionary::lookupOrDefault(const std::string& key, const std::string& def) const {
    return found(key) ? word(key) : def;
}

bool dictionary::switchOrDefault(const std::string& key, bool def) const {
    if (!found(key)) return def;
    const std::string w = word(key);
    if (w == "on" || w == "yes" || w == "true" || w == "y" || w == "1") return true;
    if (w == "off" || w == "no" || w == "false" || w == "n" || w == "0" || w == "none") return false;
    fatal("is not a switch (on/off, yes/no, true/false)", key);
}

const std::vector<NodePtr>& dictionary::list(const std::string& key) const {
    const auto& v = lookup(key);
    if (v.size() != 1 || v[0]->kind != Node::List) fatal("is not a list", key);
    return v[0]->items;
}

std::vector<double> dictionary::scalarList(const std::string& key) const {
    const auto& v = lookup(key);
    if (v.size() != 1) fatal("is not a list", key);
    return numbers(v[0], name_ + "/" + key);
}

std::vector<double> dictionary::scalarListOrDefault(const std::string& key, std::vector<double> def) const {
    return found(key) ? scalarList(key) : def;
}

std::vector<std::string> dictionary::wordList(const std::string& key) const {
    std::vector<std::string> out;
    for (const auto& it : list(key)) {
        if (it->kind != Node::Word && it->kind != Node::String) fatal("is not a list of words", key);
        out.push_back(it->text);
    }
    return out;
}

}  // namespace hs
