#include "io.hpp"

#include <math.h>
#include <stdlib.h>

#include <fstream>
#include <iostream>
#include <sstream>

namespace {

bool parse_double(const std::string &cell, double &out) {
  const char *s = cell.c_str();
  while (*s == ' ' || *s == '\t') ++s;
  if (!*s) return false;
  char *end = nullptr;
  out = strtod(s, &end);
  if (end == s) return false;
  while (*end == ' ' || *end == '\t') ++end;
  return *end == '\0';
}

}  // namespace

std::shared_ptr<Mat> CSVParser::read_csv_to_mat(std::string filename,
                                                std::vector<std::string> &column_names) {
  std::ifstream in(filename, std::ios::binary);
  if (!in.is_open()) {
    std::cout << "Could not find file '" << filename << '\'' << std::endl;
    exit(1);
  }
  std::stringstream buf;
  buf << in.rdbuf();
  const std::string text = buf.str();
  std::vector<std::vector<double>> cols;
  size_t pos = 0;
  bool first_line = true;
  while (pos < text.size()) {
    size_t eol = text.find_first_of("\r\n", pos);
    if (eol == std::string::npos) eol = text.size();
    const std::string line = text.substr(pos, eol - pos);
    pos = eol + 1;
    if (line.empty()) continue;  // "\r\n" pairs and blank lines
    std::vector<std::string> cells;
    size_t c0 = 0;
    for (;;) {
      const size_t c1 = line.find(',', c0);
      cells.push_back(line.substr(c0, c1 == std::string::npos ? std::string::npos : c1 - c0));
      if (c1 == std::string::npos) break;
      c0 = c1 + 1;
    }
    double v;
    if (first_line && !parse_double(cells[0], v)) {
      column_names = cells;
      first_line = false;
      continue;
    }
    first_line = false;
    if (cols.empty()) cols.resize(cells.size());
    for (size_t c = 0; c < cols.size(); ++c) {
      double x = 0.0;
      if (c < cells.size() && parse_double(cells[c], x)) cols[c].push_back(x);
      else cols[c].push_back(NAN);  // empty / non-numeric cell
    }
  }
  auto mat = std::make_shared<Mat>();
  mat->n_cols = cols.size();
  mat->n_rows = cols.empty() ? 0 : cols[0].size();
  mat->data.resize(mat->n_rows * mat->n_cols);
  for (size_t c = 0; c < cols.size(); ++c)
    std::copy(cols[c].begin(), cols[c].end(), mat->colptr(c));
  return mat;
}

int printSepsets(MMState *state, bool showSepsets, bool verbose) {
  int nrEdges = 0;
  const int p = (int)state->p, mc = state->maxCondSize;
  for (int i = 0; i < p; i++) {
    for (int j = i + 1; j < p; j++) {
      if (!state->adj[(size_t)i * p + j]) {
        if (showSepsets) {
          std::string sepset_string;
          for (int k = 0; k < mc; k++) {
            const int node = state->sepSets[((size_t)i * p + j) * mc + k];
            if (node == -2) {
              std::cout << "Separation from " << i << " to " << j << "\n";
              break;
            } else if (node == -1) {
              break;
            }
            sepset_string.append(std::to_string(node));
            sepset_string.append(" ");
          }
          if (!sepset_string.empty())
            std::cout << "Separation from " << i << " to " << j << " via " << sepset_string << "\n";
        }
      } else {
        nrEdges++;
      }
    }
  }
  if (verbose) std::cout << "Total number of edges: " << nrEdges << std::endl;
  return nrEdges;
}
