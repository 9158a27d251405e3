// h5io.hpp -- HDF5 dataset I/O with the behaviour of the reference's h5_helper
// (src/sim/h5_helper.h:48-129), on the HDF5 C API.
#pragma once
#include <string>
#include <vector>

namespace swhost {

enum class H5T { F32, U8, F64 };

// Reads dataset `name` converted to `type`.  When resize_allowed is false the
// element count must equal `expect` (h5_helper::read).  dims receives the shape.
bool h5_read(const std::string &file, const std::string &name, H5T type, bool resize_allowed, size_t expect,
             std::vector<unsigned char> &bytes, std::vector<size_t> &dims, std::string &err);
// Writes (replacing an existing dataset of that name; creating the file and its
// directory when missing) a dataset of the given shape (h5_helper::write).
bool h5_write(const std::string &file, const std::string &name, H5T type, const std::vector<size_t> &dims,
              const void *data, std::string &err);
bool h5_exists(const std::string &file, const std::string &name);

}  // namespace swhost
