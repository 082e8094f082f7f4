// A cursor over an OpenFOAM ASCII file (the body after its FoamFile header,
// comments blanked) and readers for its list forms, shared by the lagrangian
// cloud files and the polyMesh files.
#pragma once

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dictionary.H"

namespace hs {

// The text of an OpenFOAM ASCII file after its FoamFile header, with
// comments blanked, and a cursor over it.
class FoamText {
public:
    explicit FoamText(const std::string& path) : path_(path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw FatalError("cannot open " + path);
        std::ostringstream ss;
        ss << f.rdbuf();
        s_ = ss.str();
        blankComments();
        const size_t h = s_.find("FoamFile");
        if (h != std::string::npos) {
            const size_t close = s_.find('}', h);
            if (close == std::string::npos) fail("unterminated FoamFile header");
            i_ = close + 1;
        }
    }

    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    bool peek(char c) {
        ws();
        return i_ < s_.size() && s_[i_] == c;
    }
    void expect(char c) {
        if (!peek(c)) fail(std::string("expected '") + c + "'");
        ++i_;
    }
    double real() {
        ws();
        const char* b = s_.c_str() + i_;
        char* e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) fail("expected a number");
        i_ += (size_t)(e - b);
        return v;
    }
    long long label() {
        ws();
        const char* b = s_.c_str() + i_;
        char* e = nullptr;
        const long long v = std::strtoll(b, &e, 10);
        if (e == b) fail("expected a label");
        i_ += (size_t)(e - b);
        return v;
    }
    bool atLabel() {
        ws();
        return i_ < s_.size() && (std::isdigit((unsigned char)s_[i_]) || s_[i_] == '-');
    }
    // moves past the next occurrence of `word`
    void seek(const std::string& word) {
        const size_t k = s_.find(word, i_);
        if (k == std::string::npos) fail("no " + word);
        i_ = k + word.size();
    }
    bool next(const std::string& word) {
        ws();
        if (s_.compare(i_, word.size(), word) != 0) return false;
        i_ += word.size();
        return true;
    }
    // a word: the characters up to white space or one of ( ) { } ; "
    std::string word() {
        ws();
        const size_t b = i_;
        while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && std::string("(){};\"").find(s_[i_]) == std::string::npos)
            ++i_;
        if (i_ == b) fail("expected a word");
        return s_.substr(b, i_ - b);
    }
    // the text of an entry's value up to its ';' (which is consumed), outer
    // white space removed; parentheses nest
    std::string value() {
        ws();
        const size_t b = i_;
        int depth = 0;
        for (; i_ < s_.size(); ++i_) {
            const char c = s_[i_];
            if (c == '(') ++depth;
            else if (c == ')') --depth;
            else if (c == ';' && depth == 0) break;
            else if (c == '}' && depth == 0) fail("expected ';'");
        }
        if (i_ >= s_.size()) fail("expected ';'");
        size_t e = i_++;
        while (e > b && std::isspace((unsigned char)s_[e - 1])) --e;
        return s_.substr(b, e - b);
    }
    [[noreturn]] void fail(const std::string& what) const {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); ++k) line += s_[k] == '\n';
        throw FatalError(path_ + " line " + std::to_string(line) + ": " + what);
    }
    const std::string& path() const { return path_; }

private:
    void blankComments() {
        for (size_t k = 0; k + 1 < s_.size(); ++k) {
            if (s_[k] == '/' && s_[k + 1] == '/') {
                while (k < s_.size() && s_[k] != '\n') s_[k++] = ' ';
            } else if (s_[k] == '/' && s_[k + 1] == '*') {
                const size_t e = s_.find("*/", k + 2);
                const size_t end = e == std::string::npos ? s_.size() : e + 2;
                for (; k < end; ++k)
                    if (s_[k] != '\n') s_[k] = ' ';
                --k;
            }
        }
    }

    std::string path_, s_;
    size_t i_ = 0;
};

// An OpenFOAM list of n items, full "N ( item ... )" or uniform "N{item}":
// one(t) parses an item into a value, put(k, value) stores item k.
template <class T, class One, class Put>
int64_t readList(FoamText& t, One one, Put put) {
    const long long n = t.label();
    if (n < 0) t.fail("negative list size");
    if (t.peek('{')) {
        t.expect('{');
        const T v = one(t);
        t.expect('}');
        for (long long k = 0; k < n; ++k) put(k, v);
        return n;
    }
    t.expect('(');
    for (long long k = 0; k < n; ++k) put(k, one(t));
    t.expect(')');
    return n;
}

struct Vec3 {
    double x, y, z;
};

inline Vec3 vec3(FoamText& t) {
    t.expect('(');
    Vec3 v{ t.real(), t.real(), t.real() };
    t.expect(')');
    return v;
}

// labelList "k(l0 ... lk-1)" or uniform "k{l}"
inline std::vector<long long> labelList(FoamText& t) {
    std::vector<long long> v;
    readList<long long>(t, [](FoamText& u) { return u.label(); },
                        [&](long long k, long long x) {
                            if ((size_t)k >= v.size()) v.resize(k + 1);
                            v[k] = x;
                        });
    return v;
}

}  // namespace hs
