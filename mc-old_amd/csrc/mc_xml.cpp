// mc_xml.cpp — see mc_xml.hpp.
#include "mc_xml.hpp"

#include <cstdlib>
#include <cstring>

namespace mc {

const std::string* XmlNode::attr(const char* key) const
{
    for (const auto& a : attrs)
        if (a.first == key) return &a.second;
    return nullptr;
}

double XmlNode::as_double(const char* key, double dflt) const
{
    const std::string* v = attr(key);
    if (!v) return dflt;
    return std::strtod(v->c_str(), nullptr);
}

int XmlNode::as_int(const char* key, int dflt) const
{
    const std::string* v = attr(key);
    if (!v) return dflt;
    return (int)std::strtol(v->c_str(), nullptr, 10);
}

std::string XmlNode::value(const char* key) const
{
    const std::string* v = attr(key);
    return v ? *v : std::string();
}

const XmlNode* XmlNode::child(const char* n) const
{
    for (const auto& c : children)
        if (c.name == n) return &c;
    return nullptr;
}

std::vector<const XmlNode*> XmlNode::children_named(const char* n) const
{
    std::vector<const XmlNode*> out;
    for (const auto& c : children)
        if (c.name == n) out.push_back(&c);
    return out;
}

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;
    std::string err;

    explicit Parser(const std::string& t) : s(t) {}

    bool eof() const { return i >= s.size(); }
    void skip_ws()
    {
        while (!eof() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool starts(const char* p) const { return s.compare(i, std::strlen(p), p) == 0; }
    bool skip_until(const char* p)
    {
        const size_t k = s.find(p, i);
        if (k == std::string::npos) return false;
        i = k + std::strlen(p);
        return true;
    }
    static bool name_char(char c)
    {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
               c == '_' || c == '-' || c == '.' || c == ':';
    }
    std::string name()
    {
        const size_t b = i;
        while (!eof() && name_char(s[i])) ++i;
        return s.substr(b, i - b);
    }
    static std::string decode(const std::string& v)
    {
        std::string o;
        for (size_t k = 0; k < v.size(); ++k) {
            if (v[k] == '&') {
                const size_t e = v.find(';', k);
                if (e != std::string::npos) {
                    const std::string ent = v.substr(k + 1, e - k - 1);
                    const char* rep = ent == "lt" ? "<" : ent == "gt" ? ">" : ent == "amp" ? "&"
                                    : ent == "quot" ? "\"" : ent == "apos" ? "'" : nullptr;
                    if (rep) { o += rep; k = e; continue; }
                }
            }
            o += v[k];
        }
        return o;
    }
    // skips text, comments, processing instructions and DOCTYPE up to the next element
    // tag or closing tag; returns false at end of input
    bool next_tag()
    {
        for (;;) {
            while (!eof() && s[i] != '<') ++i;
            if (eof()) return false;
            if (starts("<!--")) { if (!skip_until("-->")) { err = "unterminated comment"; return false; } continue; }
            if (starts("<?")) { if (!skip_until("?>")) { err = "unterminated processing instruction"; return false; } continue; }
            if (starts("<![CDATA[")) { if (!skip_until("]]>")) { err = "unterminated CDATA"; return false; } continue; }
            if (starts("<!")) { if (!skip_until(">")) { err = "unterminated declaration"; return false; } continue; }
            return true;
        }
    }
    bool element(XmlNode& n)
    {
        ++i;  // '<'
        n.name = name();
        if (n.name.empty()) { err = "empty element name at offset " + std::to_string(i); return false; }
        for (;;) {
            skip_ws();
            if (eof()) { err = "unterminated tag <" + n.name; return false; }
            if (starts("/>")) { i += 2; return true; }
            if (s[i] == '>') { ++i; break; }
            std::string key = name();
            if (key.empty()) { err = "bad attribute in <" + n.name + ">"; return false; }
            skip_ws();
            if (eof() || s[i] != '=') { err = "attribute " + key + " without value in <" + n.name + ">"; return false; }
            ++i;
            skip_ws();
            if (eof() || (s[i] != '"' && s[i] != '\'')) { err = "unquoted attribute " + key; return false; }
            const char q = s[i++];
            const size_t e = s.find(q, i);
            if (e == std::string::npos) { err = "unterminated attribute " + key; return false; }
            n.attrs.emplace_back(key, decode(s.substr(i, e - i)));
            i = e + 1;
        }
        for (;;) {
            if (!next_tag()) { if (err.empty()) err = "missing </" + n.name + ">"; return false; }
            if (starts("</")) {
                i += 2;
                const std::string close = name();
                skip_ws();
                if (close != n.name || eof() || s[i] != '>') { err = "mismatched </" + close + "> for <" + n.name + ">"; return false; }
                ++i;
                return true;
            }
            n.children.emplace_back();
            if (!element(n.children.back())) return false;
        }
    }
};

}  // namespace

bool xml_parse(const std::string& text, std::vector<XmlNode>& roots, std::string& err)
{
    Parser p(text);
    roots.clear();
    while (p.next_tag()) {
        if (p.starts("</")) { err = "unexpected closing tag"; return false; }
        roots.emplace_back();
        if (!p.element(roots.back())) { err = p.err; return false; }
    }
    if (!p.err.empty()) { err = p.err; return false; }
    return true;
}

}  // namespace mc
