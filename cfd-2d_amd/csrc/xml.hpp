// Minimal XML DOM for the FVM_TVD task file.
//
// Covers what the task files use: the <?xml?> declaration, comments,
// elements with quoted attributes, text content and self-closing tags.
// Lookup semantics follow the calls the reference makes on its parser
// (src/methods/fvm_tvd.cpp:24-128): first child by name, next sibling by
// name, attribute as text/int/double, element text.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace cfd2d {

struct XmlNode {
	std::string name;
	std::string text;
	std::vector<std::pair<std::string, std::string>> attrs;
	std::vector<std::unique_ptr<XmlNode>> children;
	XmlNode* parent = nullptr;

	XmlNode* first(const char* child) const;
	XmlNode* next(const char* sibling) const;
	const char* attr(const char* key) const;  // nullptr if absent
	bool attr(const char* key, double* out) const;
	bool attr(const char* key, int* out) const;
};

class XmlDoc {
public:
	// returns false and fills error() on failure
	bool load_file(const std::string& path);
	bool parse(const std::string& text);
	const std::string& error() const { return err_; }
	XmlNode* root() const { return root_.get(); }
	XmlNode* first(const char* name) const { return root_ ? root_->first(name) : nullptr; }

private:
	std::unique_ptr<XmlNode> root_;
	std::string err_;
};

}  // namespace cfd2d
