#include "xml.hpp"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cfd2d {

XmlNode* XmlNode::first(const char* child) const
{
	for (auto& c : children)
		if (c->name == child) return c.get();
	return nullptr;
}

XmlNode* XmlNode::next(const char* sibling) const
{
	if (!parent) return nullptr;
	bool seen = false;
	for (auto& c : parent->children) {
		if (seen && c->name == sibling) return c.get();
		if (c.get() == this) seen = true;
	}
	return nullptr;
}

const char* XmlNode::attr(const char* key) const
{
	for (auto& a : attrs)
		if (a.first == key) return a.second.c_str();
	return nullptr;
}

bool XmlNode::attr(const char* key, double* out) const
{
	const char* v = attr(key);
	if (!v) return false;
	char* end = nullptr;
	double d = strtod(v, &end);
	if (end == v) return false;
	*out = d;
	return true;
}

bool XmlNode::attr(const char* key, int* out) const
{
	const char* v = attr(key);
	if (!v) return false;
	char* end = nullptr;
	long d = strtol(v, &end, 10);
	if (end == v) return false;
	*out = (int)d;
	return true;
}

namespace {

std::string decode(const std::string& s)
{
	std::string o;
	o.reserve(s.size());
	for (size_t i = 0; i < s.size(); i++) {
		if (s[i] == '&') {
			static const char* ent[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
			bool hit = false;
			for (auto& e : ent) {
				size_t n = strlen(e[0]);
				if (s.compare(i, n, e[0]) == 0) {
					o += e[1];
					i += n - 1;
					hit = true;
					break;
				}
			}
			if (!hit) o += s[i];
		} else {
			o += s[i];
		}
	}
	return o;
}

std::string trim(const std::string& s)
{
	size_t a = s.find_first_not_of(" \t\r\n");
	if (a == std::string::npos) return "";
	size_t b = s.find_last_not_of(" \t\r\n");
	return s.substr(a, b - a + 1);
}

struct Parser {
	const std::string& s;
	size_t i = 0;
	std::string err;

	explicit Parser(const std::string& t) : s(t) {}

	void skip_ws()
	{
		while (i < s.size() && isspace((unsigned char)s[i])) i++;
	}

	bool skip_misc()
	{
		for (;;) {
			skip_ws();
			if (s.compare(i, 4, "<!--") == 0) {
				size_t e = s.find("-->", i + 4);
				if (e == std::string::npos) { err = "unterminated comment"; return false; }
				i = e + 3;
			} else if (s.compare(i, 2, "<?") == 0) {
				size_t e = s.find("?>", i + 2);
				if (e == std::string::npos) { err = "unterminated declaration"; return false; }
				i = e + 2;
			} else if (s.compare(i, 2, "<!") == 0) {
				size_t e = s.find('>', i + 2);
				if (e == std::string::npos) { err = "unterminated directive"; return false; }
				i = e + 1;
			} else {
				return true;
			}
		}
	}

	std::string name()
	{
		size_t a = i;
		while (i < s.size() && (isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == '.' || s[i] == ':')) i++;
		return s.substr(a, i - a);
	}

	std::unique_ptr<XmlNode> element(XmlNode* parent)
	{
		if (i >= s.size() || s[i] != '<') { err = "expected '<'"; return nullptr; }
		i++;
		auto n = std::make_unique<XmlNode>();
		n->parent = parent;
		n->name = name();
		if (n->name.empty()) { err = "empty element name"; return nullptr; }
		for (;;) {
			skip_ws();
			if (i >= s.size()) { err = "unexpected end in tag"; return nullptr; }
			if (s[i] == '/') {
				if (s.compare(i, 2, "/>") != 0) { err = "bad '/'"; return nullptr; }
				i += 2;
				return n;
			}
			if (s[i] == '>') { i++; break; }
			std::string k = name();
			if (k.empty()) { err = "bad attribute in <" + n->name + ">"; return nullptr; }
			skip_ws();
			if (i >= s.size() || s[i] != '=') { err = "expected '=' after " + k; return nullptr; }
			i++;
			skip_ws();
			if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) { err = "expected quote"; return nullptr; }
			char q = s[i++];
			size_t e = s.find(q, i);
			if (e == std::string::npos) { err = "unterminated attribute"; return nullptr; }
			n->attrs.emplace_back(k, decode(s.substr(i, e - i)));
			i = e + 1;
		}
		std::string text;
		for (;;) {
			if (i >= s.size()) { err = "unterminated <" + n->name + ">"; return nullptr; }
			if (s.compare(i, 4, "<!--") == 0) {
				size_t e = s.find("-->", i + 4);
				if (e == std::string::npos) { err = "unterminated comment"; return nullptr; }
				i = e + 3;
				continue;
			}
			if (s.compare(i, 9, "<![CDATA[") == 0) {
				size_t e = s.find("]]>", i + 9);
				if (e == std::string::npos) { err = "unterminated CDATA"; return nullptr; }
				text += s.substr(i + 9, e - i - 9);
				i = e + 3;
				continue;
			}
			if (s.compare(i, 2, "</") == 0) {
				i += 2;
				std::string close = name();
				skip_ws();
				if (close != n->name || i >= s.size() || s[i] != '>') {
					err = "mismatched </" + close + "> for <" + n->name + ">";
					return nullptr;
				}
				i++;
				n->text = trim(decode(text));
				return n;
			}
			if (s[i] == '<') {
				auto c = element(n.get());
				if (!c) return nullptr;
				n->children.push_back(std::move(c));
				continue;
			}
			text += s[i++];
		}
	}
};

}  // namespace

bool XmlDoc::parse(const std::string& text)
{
	Parser p(text);
	root_ = std::make_unique<XmlNode>();
	root_->name = "#document";
	if (!p.skip_misc()) { err_ = p.err; return false; }
	while (p.i < text.size()) {
		auto e = p.element(root_.get());
		if (!e) { err_ = p.err; return false; }
		root_->children.push_back(std::move(e));
		if (!p.skip_misc()) { err_ = p.err; return false; }
	}
	if (root_->children.empty()) { err_ = "no root element"; return false; }
	return true;
}

bool XmlDoc::load_file(const std::string& path)
{
	std::ifstream f(path, std::ios::binary);
	if (!f) { err_ = "Failed to open file"; return false; }
	std::stringstream ss;
	ss << f.rdbuf();
	return parse(ss.str());
}

}  // namespace cfd2d
