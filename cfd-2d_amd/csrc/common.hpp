// Error bookkeeping shared by the C ABI translation units.
#pragma once

#include <string>

#include "mesh.hpp"

struct cfd2d_mesh {
	cfd2d::Mesh m;
};

namespace cfd2d {

// records msg as the calling thread's last error and returns code
int fail(int code, const std::string& msg);
void clear_error();

}  // namespace cfd2d
