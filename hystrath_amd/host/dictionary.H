// dictionary.H -- reader for the OpenFOAM dictionary files a dsmcFoam+ case
// is driven by (constant/dsmcProperties, system/boundariesDict, ...).
//
// Supported: FoamFile headers, keyword/value entries ended by ';',
// sub-dictionaries '{ }', nested lists '( )' whose items may be words,
// numbers, strings, lists or dictionaries, and // and /* */ comments.
// Errors throw hs::FatalError with the file name and line.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hs {

struct FatalError : std::runtime_error {
    explicit FatalError(const std::string& m) : std::runtime_error(m) {}
};

struct Node;
using NodePtr = std::shared_ptr<Node>;

struct Node {
    enum Kind { Word, String, List, Dict } kind = Word;
    std::string text;                  // Word / String
    std::vector<NodePtr> items;        // List
    // Dict: ordered (keyword, value tokens); a sub-dictionary is one Dict token
    std::vector<std::pair<std::string, std::vector<NodePtr>>> entries;
    int line = 0;
    bool isNumber() const;
    double number() const;             // throws if not a number
};

class dictionary {
public:
    dictionary() : node_(std::make_shared<Node>()), name_("<empty>") { node_->kind = Node::Dict; }
    dictionary(NodePtr n, std::string name) : node_(std::move(n)), name_(std::move(name)) {}

    const std::string& name() const { return name_; }
    bool found(const std::string& key) const;
    bool isDict(const std::string& key) const;
    const std::vector<NodePtr>& lookup(const std::string& key) const;   // FatalError if absent
    dictionary subDict(const std::string& key) const;
    dictionary subOrEmptyDict(const std::string& key) const;
    std::vector<std::string> toc() const;

    double scalar(const std::string& key) const;
    double lookupOrDefault(const std::string& key, double def) const;
    int label(const std::string& key) const;
    int lookupOrDefault(const std::string& key, int def) const;
    std::string word(const std::string& key) const;
    std::string lookupOrDefault(const std::string& key, const std::string& def) const;
    bool switchOrDefault(const std::string& key, bool def) const;    // on/off, yes/no, true/false
    std::vector<double> scalarList(const std::string& key) const;     // (a b c)
    std::vector<double> scalarListOrDefault(const std::string& key, std::vector<double> def) const;
    std::vector<std::string> wordList(const std::string& key) const;
    // the single List token of an entry, e.g. "dsmcPatchBoundaries ( ... );"
    const std::vector<NodePtr>& list(const std::string& key) const;

    const NodePtr& node() const { return node_; }

private:
    [[noreturn]] void fatal(const std::string& what, const std::string& key) const;
    NodePtr node_;
    std::string name_;
};

// Reads a dictionary file.
dictionary readDictionary(const std::string& path);
// Parses dictionary text (name is used in messages).
dictionary parseDictionary(const std::string& text, const std::string& name);

std::vector<double> numbers(const NodePtr& list, const std::string& what);

}  // namespace hs
