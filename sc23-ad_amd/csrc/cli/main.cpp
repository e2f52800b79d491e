// sc23ad_run: drop-in for the reference's drivers (Work/main.cpp, Work/sub.cpp)
// on one GPU, through the C ABI only.
//
//   sc23ad_run <matrix> <smoother> <solver> <m>          (Work/main.cpp:35-47)
//   sc23ad_run --sub <matrix> <smoother> <m> <bit>       (Work/sub.cpp:38-39,89)
//
// <matrix> is a MatrixMarket file (read as given, else under ../Matrix/ as the
// reference does, main.cpp:40-43) or a generated system:
//   gen:poisson2d:<n>  gen:laplace3d:<n>  gen:diffusion27:<n>:<block>:<ax>:<ay>:<az>:<seed>
//   gen:convdiff2d:<n>:<pe>:<angle>  gen:inclusions3d:<n>:<ns>:<rmin>:<rmax>:<k>:<seed>
// For a MatrixMarket file b_i = u_i in [0,1) from mt19937_64(0) with
// u = (x >> 11) 2^-53 (the reference's uniform_real_distribution is
// implementation-defined); for a generated system b = A x*, x* = 2u - 1 from
// mt19937_64(--seed, default 0).  Then LMax1Scaling (main.cpp:61).
// smoother 0 = SOR, 1 = BlockSOR(40); solver 0 FP64, 1 FP32-mixed, 2 integer.
// Options: --eps E (1e-10), --bit F (solver 2; default rule main.cpp:153-160),
// --seed S, --quiet (no per-iteration lines).
//
// stdout lines as the reference prints them (parsed by Work/Artwork/*.py):
//   # [j] <implicit relres>   per Arnoldi step (gmres.hpp:253-255)
//   [Res] : <implicit relres> on convergence (gmres.hpp:256-259)
//   [Time] : <seconds>        wall clock around the solve (timer.hpp:26-30)
//   [Iter] : <iterations>     0 when not converged (main.cpp:101)
//   # [Res] : <explicit relres> (main.cpp:176-180)
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "sc23ad.h"

namespace {

int die(const char* what) {
  std::fprintf(stderr, "sc23ad_run: %s: %s\n", what, sc23ad_last_error());
  return 1;
}

int usage() {
  std::fprintf(stderr,
               "usage: sc23ad_run <matrix> <smoother> <solver> <m> [options]\n"
               "       sc23ad_run --sub <matrix> <smoother> <m> <bit> [options]\n"
               "smoother: 0 SOR, 1 BlockSOR(40), 2 JacobiSOR (1200-row blocks)\n"
               "options: --eps E  --bit F  --seed S  --quiet\n");
  return 2;
}

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t p = 0;
  while (true) {
    size_t q = s.find(sep, p);
    out.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return out;
}

bool generated(const std::string& spec, sc23ad_csrmat** a) {
  auto f = split(spec, ':');
  if (f.size() < 3 || f[0] != "gen") return false;
  static const char* names[] = {"poisson2d", "laplace3d", "diffusion27", "convdiff2d",
                                "inclusions3d"};
  int kind = -1;
  for (int k = 0; k < 5; ++k)
    if (f[1] == names[k]) kind = k;
  if (kind < 0) return false;
  std::vector<double> prm;
  for (size_t i = 2; i < f.size(); ++i) prm.push_back(std::atof(f[i].c_str()));
  return sc23ad_csrmat_gen(kind, prm.data(), static_cast<int>(prm.size()), a) == SC23AD_OK;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  double eps = 1.0e-10;
  int bit_opt = -1;
  uint64_t seed = 0;
  bool quiet = false, sub = false;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--sub") sub = true;
    else if (a == "--quiet") quiet = true;
    else if (a == "--eps" && i + 1 < argc) eps = std::atof(argv[++i]);
    else if (a == "--bit" && i + 1 < argc) bit_opt = std::atoi(argv[++i]);
    else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
    else if (a.rfind("--", 0) == 0) return usage();
    else pos.push_back(a);
  }
  if (pos.size() != 4) return usage();
  const std::string mat = pos[0];
  const int smoother = std::atoi(pos[1].c_str());
  const int solver = sub ? SC23AD_INT : std::atoi(pos[2].c_str());
  const int m = std::atoi((sub ? pos[2] : pos[3]).c_str());
  if (sub) bit_opt = std::atoi(pos[3].c_str());
  if (smoother < 0 || smoother > 2 || solver < 0 || solver > 2 || m < 1) return usage();

  sc23ad_csrmat* A = nullptr;
  const bool gen = generated(mat, &A);
  if (!gen) {
    if (mat.rfind("gen:", 0) == 0) return die("bad generator spec");
    if (sc23ad_csrmat_read_mm(mat.c_str(), &A) != SC23AD_OK &&
        sc23ad_csrmat_read_mm(("../Matrix/" + mat).c_str(), &A) != SC23AD_OK)
      return die("reading the matrix");
  }
  int n = 0, mcols = 0, nnz = 0;
  sc23ad_csrmat_info(A, &n, &mcols, &nnz);
  std::vector<double> b(n), x(n, 0.0);
  if (gen) {
    std::vector<double> xs(n);
    if (sc23ad_xstar(n, seed, xs.data()) != SC23AD_OK) return die("x*");
    if (sc23ad_csrmat_matvec(A, xs.data(), b.data()) != SC23AD_OK) return die("b = A x*");
  } else {
    std::mt19937_64 engine(seed);
    for (int i = 0; i < n; ++i) b[i] = static_cast<double>(engine() >> 11) * 0x1.0p-53;
  }
  if (sc23ad_csrmat_lmax1_scaling(A, b.data()) != SC23AD_OK) return die("scaling");
  double nrm_b = 0;  // blas1::Nrm2 left to right (blas1.hpp:62-68)
  for (int i = 0; i < n; ++i) nrm_b += b[i] * b[i];
  nrm_b = std::sqrt(nrm_b);

  sc23ad_amg_param prm;
  sc23ad_amg_param_default(&prm, smoother);
  // The integer solver with the default bit rule needs the FP64 hierarchy too.
  const bool need_fp64 = solver == SC23AD_INT && bit_opt < 0;
  int precs[2] = {solver, SC23AD_FP64};
  sc23ad_hier* hs[2] = {nullptr, nullptr};
  const int nh = (need_fp64 ? 2 : 1);
  if (sc23ad_hier_build_multi(A, &prm, precs, nh, hs) != SC23AD_OK) return die("AMG setup");
  sc23ad_amg* M = nullptr;
  if (sc23ad_amg_create(hs[0], &M) != SC23AD_OK) return die("uploading the hierarchy");
  int bit = bit_opt;
  if (need_fp64) {
    sc23ad_amg* Md = nullptr;
    if (sc23ad_amg_create(hs[1], &Md) != SC23AD_OK) return die("uploading the FP64 hierarchy");
    if (sc23ad_default_frac_bits(Md, b.data(), nrm_b, &bit) != SC23AD_OK)
      return die("default fractional bits");
    sc23ad_amg_destroy(Md);
  }
  if (solver != SC23AD_INT) bit = 29;
  if (solver == SC23AD_INT && !quiet) std::printf("[Bit] : %d\n", bit);
  sc23ad_solver* S = nullptr;
  if (sc23ad_solver_create(solver, A, M, m, eps, bit, 1000 / m, &S) != SC23AD_OK)
    return die("creating the solver");
  sc23ad_converge cv;
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = sc23ad_solver_solve(S, b.data(), nrm_b, x.data(), 0, &cv);
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc != SC23AD_OK) return die("solve");
  if (!quiet) {
    int len = 0;
    sc23ad_solver_history(S, nullptr, 0, &len);
    std::vector<double> hist(static_cast<size_t>(len) + 1);
    sc23ad_solver_history(S, hist.data(), len, &len);
    int j = 0, restart_len = m;
    for (int k = 0; k < len; ++k) {
      std::printf("# [%d] %e\n", j + 1, hist[k]);
      if (++j == restart_len) j = 0;
    }
  }
  if (cv.is_converged) std::printf("[Res] : %e\n", cv.implicit_relres);
  std::printf("[Time] : %.15f\n", secs);
  std::printf("[Iter] : %d\n", cv.is_converged ? cv.iter : 0);
  double rel = 0;
  if (sc23ad_solver_relres(S, b.data(), nrm_b, x.data(), 0, &rel) != SC23AD_OK)
    return die("explicit residual");
  std::printf("# [Res] : %e\n", rel);
  if (sub) std::printf("\n");
  sc23ad_solver_destroy(S);
  sc23ad_amg_destroy(M);
  for (int k = 0; k < nh; ++k) sc23ad_hier_destroy(hs[k]);
  sc23ad_csrmat_destroy(A);
  return 0;
}
