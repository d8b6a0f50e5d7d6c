#include "hamx_xml.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace hamx {
namespace xml {

const Element *Element::child(const std::string &n) const {
  for (const auto &c : children)
    if (c->name == n)
      return c.get();
  return nullptr;
}

std::vector<const Element *> Element::children_named(const std::string &n) const {
  std::vector<const Element *> out;
  for (const auto &c : children)
    if (c->name == n)
      out.push_back(c.get());
  return out;
}

const std::string *Element::attribute(const std::string &n) const {
  for (const auto &a : attributes)
    if (a.first == n)
      return &a.second;
  return nullptr;
}

double Element::attr_double(const std::string &n) const {
  const std::string *v = attribute(n);
  return v ? std::strtod(v->c_str(), nullptr) : 0.0;
}

unsigned long long Element::attr_uint(const std::string &n) const {
  const std::string *v = attribute(n);
  return v ? std::strtoull(v->c_str(), nullptr, 10) : 0ull;
}

bool Element::attr_bool(const std::string &n) const {
  const std::string *v = attribute(n);
  if (!v)
    return false;
  std::string s;
  for (char c : *v)
    if (!std::isspace(static_cast<unsigned char>(c)))
      s += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  if (s == "true" || s == "1")
    return true;
  if (s == "false" || s == "0")
    return false;
  return std::strtol(s.c_str(), nullptr, 10) != 0;
}

std::string Element::attr_string(const std::string &n) const {
  const std::string *v = attribute(n);
  if (!v)
    throw std::runtime_error("missing attribute '" + n + "' on <" + name + ">");
  return *v;
}

namespace {

struct Reader {
  const std::string &s;
  size_t i = 0;
  explicit Reader(const std::string &text) : s(text) {}

  [[noreturn]] void fail(const std::string &what) const {
    throw std::runtime_error("XML parse error at offset " + std::to_string(i) + ": " + what);
  }
  bool at(const char *lit) const { return s.compare(i, std::char_traits<char>::length(lit), lit) == 0; }
  void skip_ws() {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i])))
      ++i;
  }
  void skip_past(const char *lit) {
    const size_t p = s.find(lit, i);
    if (p == std::string::npos)
      fail(std::string("unterminated construct, expected ") + lit);
    i = p + std::char_traits<char>::length(lit);
  }
  // comments, declarations, processing instructions, CDATA between elements
  void skip_misc() {
    for (;;) {
      skip_ws();
      if (at("<!--"))
        skip_past("-->");
      else if (at("<?"))
        skip_past("?>");
      else if (at("<![CDATA["))
        skip_past("]]>");
      else if (at("<!"))
        skip_past(">");
      else
        return;
    }
  }
  std::string name() {
    const size_t b = i;
    while (i < s.size() && (std::isalnum(static_cast<unsigned char>(s[i])) || s[i] == '_' ||
                            s[i] == '-' || s[i] == '.' || s[i] == ':'))
      ++i;
    if (i == b)
      fail("expected a name");
    return s.substr(b, i - b);
  }
  static std::string unescape(const std::string &v) {
    std::string o;
    for (size_t k = 0; k < v.size(); ++k) {
      if (v[k] != '&') {
        o += v[k];
        continue;
      }
      const size_t e = v.find(';', k);
      if (e == std::string::npos) {
        o += v[k];
        continue;
      }
      const std::string ent = v.substr(k + 1, e - k - 1);
      if (ent == "amp") o += '&';
      else if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else if (!ent.empty() && ent[0] == '#')
        o += static_cast<char>(ent.size() > 1 && ent[1] == 'x'
                                   ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                   : std::strtol(ent.c_str() + 1, nullptr, 10));
      else
        o += "&" + ent + ";";
      k = e;
    }
    return o;
  }
  std::unique_ptr<Element> element() {
    if (i >= s.size() || s[i] != '<')
      fail("expected '<'");
    ++i;
    auto e = std::make_unique<Element>();
    e->name = name();
    for (;;) {
      skip_ws();
      if (i >= s.size())
        fail("unterminated start tag <" + e->name + ">");
      if (at("/>")) {
        i += 2;
        return e;
      }
      if (s[i] == '>') {
        ++i;
        break;
      }
      std::string an = name();
      skip_ws();
      if (i >= s.size() || s[i] != '=')
        fail("expected '=' after attribute " + an);
      ++i;
      skip_ws();
      if (i >= s.size() || (s[i] != '"' && s[i] != '\''))
        fail("expected a quoted value for attribute " + an);
      const char q = s[i++];
      const size_t end = s.find(q, i);
      if (end == std::string::npos)
        fail("unterminated attribute value");
      e->attributes.emplace_back(an, unescape(s.substr(i, end - i)));
      i = end + 1;
    }
    // content
    for (;;) {
      // text up to the next tag is ignored
      while (i < s.size() && s[i] != '<')
        ++i;
      if (i >= s.size())
        fail("missing end tag </" + e->name + ">");
      if (at("</")) {
        i += 2;
        const std::string en = name();
        if (en != e->name)
          fail("mismatched end tag </" + en + "> for <" + e->name + ">");
        skip_ws();
        if (i >= s.size() || s[i] != '>')
          fail("expected '>'");
        ++i;
        return e;
      }
      if (at("<!--") || at("<?") || at("<!")) {
        skip_misc();
        continue;
      }
      e->children.push_back(element());
    }
  }
};

} // namespace

Document parse(const std::string &text) {
  Reader r(text);
  r.skip_misc();
  Document d;
  d.root = r.element();
  r.skip_misc();
  return d;
}

Document load(const std::string &filename) {
  std::ifstream f(filename);
  if (!f)
    throw std::runtime_error("cannot open parameter file " + filename);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

const Element *trace(const Element *e, const std::vector<std::string> &chain) {
  for (const auto &k : chain) {
    const Element *c = e ? e->child(k) : nullptr;
    if (!c)
      throw std::runtime_error("missing XML element <" + k + ">");
    e = c;
  }
  return e;
}

} // namespace xml
} // namespace hamx
