// MatrixMarket input with the semantics of io::ReadMM(path, is_expanded = true,
// has_zeros = false) (Library/io/mm.hpp:24-98): coordinate real/integer,
// general or symmetric; explicit zeros dropped; entries sorted by row, then by
// column within each row, with the reference's Hoare sort; a symmetric file
// (lower triangle stored) is expanded to both triangles.
#pragma once

#include <string>

#include "csr.hpp"

namespace sc23ad {

bool read_matrix_market(const std::string& path, HostCsr& out,
                        std::string& err);

}  // namespace sc23ad
