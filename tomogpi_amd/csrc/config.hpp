// config.hpp - reference-compatible config-file lookup (host only).
#pragma once

#include <string>

#include "tomogpi.h"

namespace tomo {

// First token equal to `key` in the file, value = next token on that line
// (C/Config/Config.cpp:34-180).  TOMO_ERR_CONFIG if the file or key is missing.
int config_lookup(const std::string& path, const std::string& key, std::string* value);
int config_int(const std::string& path, const std::string& key, int* out);
int config_double(const std::string& path, const std::string& key, double* out);
std::string join_path(const std::string& dir, const std::string& file);
int read_config_ct(const std::string& workdir, tomo_geom_desc* d);
size_t dtype_size(int dtype);

}  // namespace tomo
