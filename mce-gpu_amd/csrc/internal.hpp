// Internal declarations shared by the host translation units of libmce.so.
#pragma once
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "../../include/mce.h"

namespace mce {

// Records the detail of a failure for mce_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Builds a simple symmetric CSR from an edge list: both directions, no self
// loops, rows sorted and deduplicated. n == 0 means max id + 1.
int build_csr(uint32_t n, const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst,
              mce_csr *out);

// Allocates the arrays of an mce_csr (row_ptr zeroed).
int alloc_csr(uint32_t n, uint64_t m, mce_csr *out);

}  // namespace mce
