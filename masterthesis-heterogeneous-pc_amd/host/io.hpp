// Input and output formats: the observation / correlation CSV reader
// (src/util/csv_parser.{hpp,cpp}) and the sepset printout
// (src/util/sepset_util.hpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "state.hpp"

// Column-major n_rows x n_cols matrix (the reference's arma::mat): column v
// holds variable v, so the storage is variable-major [p][n], the layout
// gpuPMCC takes (corOwn.cu:10).
struct Mat {
  size_t n_rows = 0, n_cols = 0;
  std::vector<double> data;
  const double *begin() const { return data.data(); }
  double *colptr(size_t c) { return data.data() + c * n_rows; }
};

struct CSVParser {
  // csv_parser.cpp:20-77: ',' separates cells, '\r' and '\n' end rows, and a
  // first cell that is not a number marks a header line (its cells become
  // column_names).  Exits with "Could not find file" like the reference when
  // the file cannot be opened.
  static std::shared_ptr<Mat> read_csv_to_mat(std::string filename,
                                              std::vector<std::string> &column_names);
};

// sepset_util.hpp:9-60: counts the edges left (i < j) and, when showSepsets,
// prints "Separation from i to j" (level 0) or "... via S" for every removed
// pair.  Prints the total under verbose.
int printSepsets(MMState *state, bool showSepsets, bool verbose);
