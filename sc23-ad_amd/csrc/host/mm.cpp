#include "mm.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <utility>

namespace sc23ad {

namespace {

// helper::matrix::expand(..., "L") (Library/helper/helper_matrix.hpp:730-766):
// row i = stored row i followed by column i of the stored triangle minus its
// first (diagonal) entry.
HostCsr expand_lower(const HostCsr& l) {
  const int n = l.n;
  HostCsr t;  // transpose of l
  t.reset(n, n);
  for (int p = 0; p < l.nnz(); ++p) ++t.rptr[l.cind[p] + 1];
  for (int i = 0; i < n; ++i) t.rptr[i + 1] += t.rptr[i];
  t.cind.resize(l.nnz());
  t.val.resize(l.nnz());
  std::vector<int> fill(t.rptr.begin(), t.rptr.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int p = l.rptr[i]; p < l.rptr[i + 1]; ++p) {
      const int d = fill[l.cind[p]]++;
      t.cind[d] = i;
      t.val[d] = l.val[p];
    }
  HostCsr o;
  o.reset(n, n);
  for (int i = 0; i < n; ++i) {
    for (int p = l.rptr[i]; p < l.rptr[i + 1]; ++p) {
      o.cind.push_back(l.cind[p]);
      o.val.push_back(l.val[p]);
    }
    for (int p = t.rptr[i] + 1; p < t.rptr[i + 1]; ++p) {
      o.cind.push_back(t.cind[p]);
      o.val.push_back(t.val[p]);
    }
    o.rptr[i + 1] = static_cast<int>(o.cind.size());
  }
  return o;
}

}  // namespace

bool read_matrix_market(const std::string& path, HostCsr& out,
                        std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open " + path;
    return false;
  }
  std::string line;
  if (!std::getline(in, line)) {
    err = "empty file";
    return false;
  }
  std::istringstream hs(line);
  std::string banner, object, format, field, symmetry;
  hs >> banner >> object >> format >> field >> symmetry;
  if (banner != "%%MatrixMarket" || object != "matrix" || format != "coordinate") {
    err = "only '%%MatrixMarket matrix coordinate' files are supported";
    return false;
  }
  if (field != "real" && field != "integer") {
    err = field + " is not supported";
    return false;
  }
  bool symmetric;
  if (symmetry == "general") {
    symmetric = false;
  } else if (symmetry == "symmetric") {
    symmetric = true;
  } else {
    err = symmetry + " is not supported";
    return false;
  }
  while (std::getline(in, line))
    if (!line.empty() && line[0] != '%') break;
  long long nr = 0, nc = 0, pe = 0;
  {
    std::istringstream ss(line);
    if (!(ss >> nr >> nc >> pe) || nr < 0 || nc < 0 || pe < 0 ||
        nr > 0x7fffffff || nc > 0x7fffffff || pe > 0x7fffffff) {
      err = "bad size line";
      return false;
    }
  }
  const int n = static_cast<int>(nr), m = static_cast<int>(nc);
  std::vector<int> row, col;
  std::vector<double> val;
  row.reserve(pe);
  col.reserve(pe);
  val.reserve(pe);
  for (long long k = 0; k < pe; ++k) {
    long long r, c;
    double v;
    if (!(in >> r >> c >> v)) {
      err = "truncated entry list";
      return false;
    }
    if (r < 1 || r > n || c < 1 || c > m) {
      err = "entry index out of range";
      return false;
    }
    if (v == 0) continue;
    row.push_back(static_cast<int>(r - 1));
    col.push_back(static_cast<int>(c - 1));
    val.push_back(v);
  }
  const int nnz = static_cast<int>(row.size());
  HostCsr a;
  a.reset(n, m);
  if (nnz > 0)
    hoare_sort_with(0, nnz - 1, true, row.data(), [&](int x, int y) {
      std::swap(col[x], col[y]);
      std::swap(val[x], val[y]);
    });
  int cnt = 0;
  for (int i = 0; i < n; ++i) {
    int num = 0;
    while (cnt < nnz && row[cnt] == i) {
      ++num;
      ++cnt;
    }
    a.rptr[i + 1] = a.rptr[i] + num;
    if (num > 1)
      hoare_sort_with(a.rptr[i], a.rptr[i + 1] - 1, true, col.data(),
                      [&](int x, int y) { std::swap(val[x], val[y]); });
  }
  a.cind = std::move(col);
  a.val = std::move(val);
  out = symmetric ? expand_lower(a) : std::move(a);
  return true;
}

}  // namespace sc23ad
