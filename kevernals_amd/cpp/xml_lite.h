// A small XML reader for acquisition-geometry files: elements, text,
// CDATA, the five predefined entities and numeric character references.
// Declarations, processing instructions, comments and DOCTYPE blocks are
// skipped; attributes are read past and ignored.  Whitespace between markup
// does not form nodes, as in the reader the reference uses
// (commons/tinyXML/tinyxml2.cpp:718-770 skips it before identifying a node);
// a text node keeps its surrounding whitespace.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace xml_lite {

struct Node {
    bool is_text = false;
    std::string name;  // element name
    std::string text;  // text node content
    std::vector<std::unique_ptr<Node>> children;

    const Node *first_child_element(const std::string &n) const
    {
        for (const auto &c : children)
            if (!c->is_text && c->name == n)
                return c.get();
        return nullptr;
    }
    // Text of the element when its first child is a text node, else null
    // (the GetText() convention).
    const std::string *get_text() const
    {
        return (!children.empty() && children.front()->is_text) ? &children.front()->text : nullptr;
    }
    std::vector<const Node *> child_elements() const
    {
        std::vector<const Node *> out;
        for (const auto &c : children)
            if (!c->is_text)
                out.push_back(c.get());
        return out;
    }
};

class Parser {
public:
    explicit Parser(const std::string &s) : s_(s) {}

    // Parses the document; returns the root element or null (error()).
    std::unique_ptr<Node> parse()
    {
        if (s_.compare(0, 3, "\xEF\xBB\xBF") == 0)
            i_ = 3;
        std::unique_ptr<Node> root;
        while (skip_ws(), i_ < s_.size()) {
            if (skip_misc()) {
                if (!err_.empty())
                    return nullptr;
                continue;
            }
            if (s_[i_] != '<') {
                err_ = "text outside the root element";
                return nullptr;
            }
            if (root) {
                err_ = "more than one root element";
                return nullptr;
            }
            root = element();
            if (!root)
                return nullptr;
        }
        if (!root && err_.empty())
            err_ = "no root element";
        return root;
    }
    const std::string &error() const { return err_; }

private:
    static bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
    static bool is_name(char c)
    {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_' || c == '-' ||
               c == '.' || c == ':' || (unsigned char)c >= 0x80;
    }
    void skip_ws()
    {
        while (i_ < s_.size() && is_ws(s_[i_]))
            ++i_;
    }
    bool starts(const char *p) const { return s_.compare(i_, std::char_traits<char>::length(p), p) == 0; }
    bool skip_past(const char *end)
    {
        const size_t k = s_.find(end, i_);
        if (k == std::string::npos) {
            err_ = std::string("unterminated markup, missing ") + end;
            return false;
        }
        i_ = k + std::char_traits<char>::length(end);
        return true;
    }
    // Declarations, processing instructions, comments, DOCTYPE: skipped.
    // Returns true when something was skipped (or an error was set).
    bool skip_misc()
    {
        if (starts("<?")) {
            skip_past("?>");
            return true;
        }
        if (starts("<!--")) {
            skip_past("-->");
            return true;
        }
        if (starts("<!") && !starts("<![CDATA[")) {
            int depth = 0;
            for (; i_ < s_.size(); ++i_) {
                if (s_[i_] == '[')
                    ++depth;
                else if (s_[i_] == ']')
                    --depth;
                else if (s_[i_] == '>' && depth <= 0) {
                    ++i_;
                    return true;
                }
            }
            err_ = "unterminated declaration";
            return true;
        }
        return false;
    }
    std::string name()
    {
        const size_t b = i_;
        while (i_ < s_.size() && is_name(s_[i_]))
            ++i_;
        return s_.substr(b, i_ - b);
    }
    static void put_utf8(std::string &out, uint32_t cp)
    {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    // Text up to the next '<', entities decoded.
    bool text(std::string &out)
    {
        while (i_ < s_.size() && s_[i_] != '<') {
            if (s_[i_] != '&') {
                out += s_[i_++];
                continue;
            }
            const size_t semi = s_.find(';', i_);
            if (semi == std::string::npos || semi - i_ > 12) {
                out += s_[i_++];  // a lone '&' is kept as it is
                continue;
            }
            const std::string ent = s_.substr(i_ + 1, semi - i_ - 1);
            if (ent == "lt")
                out += '<';
            else if (ent == "gt")
                out += '>';
            else if (ent == "amp")
                out += '&';
            else if (ent == "quot")
                out += '"';
            else if (ent == "apos")
                out += '\'';
            else if (ent.size() > 1 && ent[0] == '#') {
                const bool hex = ent[1] == 'x' || ent[1] == 'X';
                char *end = nullptr;
                const std::string digits = ent.substr(hex ? 2 : 1);
                const unsigned long cp = std::strtoul(digits.c_str(), &end, hex ? 16 : 10);
                if (digits.empty() || *end != '\0' || cp > 0x10FFFF) {
                    err_ = "bad character reference &" + ent + ";";
                    return false;
                }
                put_utf8(out, (uint32_t)cp);
            } else {
                out.append(s_, i_, semi - i_ + 1);  // unknown entity: kept verbatim
            }
            i_ = semi + 1;
        }
        return true;
    }
    std::unique_ptr<Node> element()
    {
        ++i_;  // '<'
        auto node = std::make_unique<Node>();
        node->name = name();
        if (node->name.empty()) {
            err_ = "element without a name";
            return nullptr;
        }
        // attributes: read past (quoted values may hold '>')
        for (;;) {
            skip_ws();
            if (i_ >= s_.size()) {
                err_ = "unterminated start tag <" + node->name;
                return nullptr;
            }
            if (starts("/>")) {
                i_ += 2;
                return node;
            }
            if (s_[i_] == '>') {
                ++i_;
                break;
            }
            if (s_[i_] == '"' || s_[i_] == '\'') {
                const char q = s_[i_++];
                const size_t k = s_.find(q, i_);
                if (k == std::string::npos) {
                    err_ = "unterminated attribute value";
                    return nullptr;
                }
                i_ = k + 1;
            } else {
                ++i_;
            }
        }
        // content
        for (;;) {
            const size_t b = i_;
            skip_ws();
            if (i_ >= s_.size()) {
                err_ = "unterminated element <" + node->name + ">";
                return nullptr;
            }
            if (s_[i_] != '<') {
                i_ = b;  // the text keeps its leading whitespace
                auto t = std::make_unique<Node>();
                t->is_text = true;
                if (!text(t->text))
                    return nullptr;
                node->children.push_back(std::move(t));
                continue;
            }
            if (starts("</")) {
                i_ += 2;
                const std::string close = name();
                skip_ws();
                if (close != node->name || i_ >= s_.size() || s_[i_] != '>') {
                    err_ = "mismatched end tag </" + close + "> for <" + node->name + ">";
                    return nullptr;
                }
                ++i_;
                return node;
            }
            if (starts("<![CDATA[")) {
                i_ += 9;
                const size_t k = s_.find("]]>", i_);
                if (k == std::string::npos) {
                    err_ = "unterminated CDATA section";
                    return nullptr;
                }
                auto t = std::make_unique<Node>();
                t->is_text = true;
                t->text = s_.substr(i_, k - i_);
                i_ = k + 3;
                node->children.push_back(std::move(t));
                continue;
            }
            if (skip_misc()) {
                if (!err_.empty())
                    return nullptr;
                continue;
            }
            auto child = element();
            if (!child)
                return nullptr;
            node->children.push_back(std::move(child));
        }
    }

    const std::string &s_;
    size_t i_ = 0;
    std::string err_;
};

inline std::unique_ptr<Node> load_file(const std::string &path, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return nullptr;
    }
    std::ostringstream ss;
    ss << f.rdbuf();
    const std::string doc = ss.str();
    Parser p(doc);
    auto root = p.parse();
    if (!root)
        err = p.error();
    return root;
}

}  // namespace xml_lite
