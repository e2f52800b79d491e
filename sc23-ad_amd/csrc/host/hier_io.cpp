#include "hier_io.hpp"

#include <cstdio>

namespace sc23ad {

namespace {

struct Writer {
  FILE* f;
  bool ok = true;
  void raw(const void* p, size_t bytes) {
    if (ok && bytes && fwrite(p, 1, bytes, f) != bytes) ok = false;
  }
  void i32(int v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
};

template <typename Mat>
void put_matrix(Writer& w, const Mat& a, int bit, int dbit) {
  w.i32(a.n);
  w.i32(a.m);
  w.i32(a.nnz());
  w.i32(bit);
  w.i32(dbit);
  w.raw(a.rptr.data(), sizeof(int) * a.rptr.size());
  w.raw(a.cind.data(), sizeof(int) * a.cind.size());
  w.raw(a.val.data(), sizeof(a.val[0]) * a.val.size());
}

void put_int(Writer& w, const IntCsr& a) { put_matrix(w, a, a.bit, a.dbit); }
template <typename T>
void put_fp(Writer& w, const FpCsr<T>& a) {
  put_matrix(w, a, 0, -1);
}

template <typename H>
void put_header(Writer& w, const H& h, int precision) {
  w.raw("SC23AMG1", 8);
  w.i32(precision);
  w.i32(h.depth);
  w.i32(static_cast<int>(h.prm.smoother));
  w.i32(h.prm.iter);
  w.i32(h.prm.b_num);
  w.i32(h.prm.max_levels);
  w.i32(h.prm.coarse_n);
  w.i32(h.n_coarse);
  w.f64(h.prm.omega);
  w.f64(h.prm.theta);
}

}  // namespace

bool write_hierarchy(const std::string& path, const IntHierarchy& h) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  Writer w{f};
  put_header(w, h, 2);
  for (const IntLevel& L : h.lv) {
    put_int(w, L.a);
    put_int(w, L.ld);
    put_int(w, L.u);
    put_int(w, L.r);
    put_int(w, L.p);
  }
  put_int(w, h.lc);
  put_int(w, h.uc);
  return fclose(f) == 0 && w.ok;
}

template <typename T>
static bool write_fp(const std::string& path, const FpHierarchy<T>& h,
                     int precision) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  Writer w{f};
  put_header(w, h, precision);
  for (const FpLevel<T>& L : h.lv) {
    put_fp(w, L.a);
    put_fp(w, L.ld);
    put_fp(w, L.u);
    put_fp(w, L.r);
    put_fp(w, L.p);
  }
  put_fp(w, h.lc);
  put_fp(w, h.uc);
  return fclose(f) == 0 && w.ok;
}

bool write_hierarchy(const std::string& path, const FpHierarchy<double>& h) {
  return write_fp(path, h, 0);
}
bool write_hierarchy(const std::string& path, const FpHierarchy<float>& h) {
  return write_fp(path, h, 1);
}

}  // namespace sc23ad
