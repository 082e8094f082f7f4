// dictCheck -- parses OpenFOAM dictionary files with the host reader and
// prints each file's top-level keywords (or the error); with -case it also
// builds the engine configuration of a case without creating an engine.
#include <cstring>
#include <iostream>

#include "dictionary.H"

static void dump(const hs::NodePtr& n, int depth) {
    std::string pad(depth * 2, ' ');
    for (const auto& e : n->entries) {
        std::cout << pad << e.first;
        for (const auto& v : e.second) {
            if (v->kind == hs::Node::Dict) {
                std::cout << " {\n";
                dump(v, depth + 1);
                std::cout << pad << "}";
            } else if (v->kind == hs::Node::List) {
                std::cout << " (" << v->items.size() << " items)";
            } else {
                std::cout << " " << v->text;
            }
        }
        std::cout << "\n";
    }
}

int main(int argc, char** argv) {
    int rc = 0;
    for (int i = 1; i < argc; ++i) {
        try {
            hs::dictionary d = hs::readDictionary(argv[i]);
            std::cout << "== " << argv[i] << "\n";
            dump(d.node(), 0);
        } catch (const hs::FatalError& e) {
            std::cout << "!! " << argv[i] << ": " << e.what() << "\n";
            rc = 1;
        }
    }
    return rc;
}
