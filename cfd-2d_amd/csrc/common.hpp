// Error bookkeeping shared by the C ABI translation units.
#pragma once

#include <string>

#include "mesh.hpp"

namespace cfd2d {
// a process-unique number (never 0), for telling meshes apart
unsigned long long next_mesh_id();
}  // namespace cfd2d

struct cfd2d_mesh {
	cfd2d::Mesh m;
	// changes whenever the geometry does (a new handle, a periodic pairing), so
	// device copies of the geometry can be kept per id
	unsigned long long id = cfd2d::next_mesh_id();
};

namespace cfd2d {

// records msg as the calling thread's last error and returns code
int fail(int code, const std::string& msg);
void clear_error();

}  // namespace cfd2d
