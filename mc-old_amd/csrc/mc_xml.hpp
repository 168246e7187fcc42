// mc_xml.hpp — minimal XML reader for the reference's input decks.
//
// The reference parses decks with vendored pugixml 1.8 (setup.cpp:34-35).
// Its decks are attribute-only element trees with several top-level
// elements and no single root, so this reader keeps exactly that: elements,
// attributes, children; prolog, comments, DOCTYPE and text are skipped.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace mc {

struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;

    const std::string* attr(const char* key) const;
    bool has(const char* key) const { return attr(key) != nullptr; }
    // pugixml-style conversions: 0 when absent or unparsable
    double as_double(const char* key, double dflt = 0.0) const;
    int as_int(const char* key, int dflt = 0) const;
    std::string value(const char* key) const;
    const XmlNode* child(const char* name) const;
    std::vector<const XmlNode*> children_named(const char* name) const;
};

// Parses text into its top-level elements.  Returns false and sets err on
// malformed input.
bool xml_parse(const std::string& text, std::vector<XmlNode>& roots, std::string& err);

}  // namespace mc
