// Runs the spams.h-compatible templates of include/spams_amd.h on the GPU the
// way a C++ caller of the reference would, on inputs written by
// tests/test_gpu_dropin.py, and writes every result back for the test to
// check against the oracle.
//
//   dropin_run <dir>   reads <dir>/dims.txt ("m n p"), X.bin (m x n), D.bin
//                      (m x p), doubles column-major; writes <dir>/<name>.bin
#include <cstdio>
#include <string>
#include <vector>

#include "spams_amd.h"

static std::string g_dir;

static bool read_doubles(const char* name, double* dst, size_t n) {
  FILE* f = std::fopen((g_dir + "/" + name).c_str(), "rb");
  if (!f) return false;
  const size_t got = std::fread(dst, sizeof(double), n, f);
  std::fclose(f);
  return got == n;
}

static void write_raw(const std::string& name, const void* p, size_t bytes) {
  FILE* f = std::fopen((g_dir + "/" + name + ".bin").c_str(), "wb");
  if (!f) return;
  std::fwrite(p, 1, bytes, f);
  std::fclose(f);
}

// CSC as: n, nnz (int64), pB[n+1] (int32), r[nnz] (int32), v[nnz] (double)
static void write_csc(const std::string& name, const SpMatrix<double>& a) {
  const int n = a.n();
  const long long nnz = n > 0 ? a.pB()[n] : 0;
  std::vector<char> buf;
  auto put = [&](const void* p, size_t b) {
    const char* c = static_cast<const char*>(p);
    buf.insert(buf.end(), c, c + b);
  };
  put(&n, sizeof(int));
  put(&nnz, sizeof(long long));
  put(a.pB(), sizeof(int) * (n + 1));
  put(a.r(), sizeof(int) * nnz);
  put(a.v(), sizeof(double) * nnz);
  write_raw(name, buf.data(), buf.size());
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  g_dir = argv[1];
  int m = 0, n = 0, p = 0;
  {
    FILE* f = std::fopen((g_dir + "/dims.txt").c_str(), "r");
    if (!f || std::fscanf(f, "%d %d %d", &m, &n, &p) != 3) return 3;
    std::fclose(f);
  }
  Matrix<double> X(m, n), D(m, p);
  if (!read_doubles("X.bin", X.rawX(), (size_t)m * n) ||
      !read_doubles("D.bin", D.rawX(), (size_t)m * p))
    return 4;
  try {
    // _lassoD (cpp_library/spams.h:107-143)
    Matrix<double>* path = nullptr;
    SpMatrix<double>* a = _lassoD<double>(&X, &D, &path, false, -1, 0.15, 0.0, PENALTY, false,
                                          false, -1, -1, false, false);
    write_csc("lassoD", *a);
    delete a;
    // decomp-level lasso with the least-squares refit (decomp/decomp.h:111-116)
    const int L = m < p ? m : p;
    SpMatrix<double> b;
    lasso(X, D, b, L, 0.15, 0.0, PENALTY, false, true);
    write_csc("lasso_ols", b);
    // Gram-given lasso2 (decomp/decomp.h:131-135): G = D'D + 1e-10 I, DtX = D'X
    Matrix<double> G(p, p), DtX(p, n);
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < p; ++j) {
        double s = 0;
        for (int r = 0; r < m; ++r) s += D(r, i) * D(r, j);
        G(i, j) = s + (i == j ? 1e-10 : 0.0);
      }
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int r = 0; r < m; ++r) s += D(r, i) * X(r, j);
        DtX(i, j) = s;
      }
    SpMatrix<double> c;
    lasso2(X, G, DtX, c, L, 0.15);
    write_csc("lasso2_gram", c);
    // _omp with a scalar L (cpp_library/spams.h:250-293)
    Vector<int> vL(1);
    vL[0] = 8;
    Vector<double> ve(1), vl(1);
    Matrix<double>* opath = nullptr;
    SpMatrix<double>* o = _omp<double>(&X, &D, &opath, false, true, &vL, false, &ve, false, &vl, -1);
    write_csc("omp", *o);
    delete o;
    // _alltrainDL from the given dictionary, with the model (cpp_library/spams.h:1153-1340)
    Matrix<double> mA, mB;
    Matrix<double>* A = nullptr;
    Matrix<double>* B = nullptr;
    Vector<int>* it = nullptr;
    Vector<double> eta;
    SpMatrix<bool> g, gv;
    Vector<int> own, nown;
    char regul[] = "none";
    char logname[] = "";
    Matrix<double>* Dl = _alltrainDL<double>(
        &X, false, &A, &B, &it, true, &mA, &mB, 0, &D, &eta, &g, &gv, &own, &nown, 1, 1e-6, true,
        false, 64, p, 0.15, 1e-9, 0., 5, 1e-5, PENALTY, regul, false, false, false, L2, false,
        true, false, 0., 0., 1.0, 1, false, 0, false, false, logname);
    write_raw("traindl_D", Dl->rawX(), sizeof(double) * (size_t)m * p);
    write_raw("traindl_A", A->rawX(), sizeof(double) * (size_t)p * p);
    write_raw("traindl_iter", it->rawX(), sizeof(int));
    delete Dl;
    delete A;
    delete B;
    delete it;
    // _lassoWeighted / _lassoMask / _ompMask (cpp_library/spams.h:190-249, :295-343)
    Matrix<double> Wt(p, n);
    std::vector<double> braw((size_t)m * n);
    if (!read_doubles("W.bin", Wt.rawX(), (size_t)p * n) ||
        !read_doubles("B.bin", braw.data(), braw.size()))
      return 6;
    Matrix<bool> Bm(m, n);
    for (size_t e = 0; e < braw.size(); ++e) Bm.rawX()[e] = braw[e] != 0.0;
    SpMatrix<double>* w = _lassoWeighted<double>(&X, &D, &Wt, -1, 0.1, PENALTY, false, -1, false);
    write_csc("lasso_weighted", *w);
    delete w;
    SpMatrix<double>* mk = _lassoMask<double>(&X, &D, &Bm, -1, 0.1, 0.0, PENALTY, false, -1, false);
    write_csc("lasso_mask", *mk);
    delete mk;
    Matrix<double>* mpath = nullptr;
    SpMatrix<double>* om = _ompMask<double>(&X, &D, &Bm, &mpath, false, true, &vL, false, &ve,
                                            false, &vl, -1);
    write_csc("omp_mask", *om);
    delete om;
    // _fistaFlat, l1, fixed step (cpp_library/spams.h:460-608); W written in place
    double L0 = 0;
    if (!read_doubles("L0.bin", &L0, 1)) return 7;
    Matrix<double> W0(p, n), Wf(p, n);
    W0.setZeros();
    Vector<int> groups;
    char rl1[] = "l1";
    char lsq[] = "square";
    char lname[] = "";
    Matrix<double>* info = _fistaFlat<double>(
        &X, &D, &W0, &Wf, &groups, 1, 60, L0, true, 1.5, 0.05, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0,
        100, 1000, false, false, false, false, false, rl1, lsq, false, false, false, false,
        false, false, lname, false, nullptr, 1, true, false, 0);
    write_raw("fista_W", Wf.rawX(), sizeof(double) * (size_t)p * n);
    delete info;
    // the same call with the dictionary as an SpMatrix (AbstractMatrixB),
    // and with compute_gram and backtracking (prox/fista.h:364-413)
    {
      std::vector<int> pB(p + 1), rr;
      std::vector<double> vv;
      for (int j = 0; j < p; ++j) {
        pB[j] = (int)rr.size();
        for (int i = 0; i < m; ++i)
          if (D(i, j) != 0.0) {
            rr.push_back(i);
            vv.push_back(D(i, j));
          }
      }
      pB[p] = (int)rr.size();
      SpMatrix<double> Ds(vv.data(), rr.data(), pB.data(), pB.data() + 1, m, p, (int)rr.size());
      Matrix<double> Ws(p, n), Wg(p, n);
      info = _fistaFlat<double>(
          &X, &Ds, &W0, &Ws, &groups, 1, 60, L0, true, 1.5, 0.05, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0,
          0.0, 100, 1000, false, false, false, false, false, rl1, lsq, false, false, false,
          false, false, false, lname, false, nullptr, 1, true, false, 0);
      write_raw("fista_sparseD_W", Ws.rawX(), sizeof(double) * (size_t)p * n);
      delete info;
      info = _fistaFlat<double>(
          &X, &D, &W0, &Wg, &groups, 1, 60, 0.1, false, 1.5, 0.05, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0,
          0.0, 100, 1000, true, false, false, false, false, rl1, lsq, false, false, false,
          false, false, false, lname, false, nullptr, 1, true, false, 0);
      write_raw("fista_gram_W", Wg.rawX(), sizeof(double) * (size_t)p * n);
      delete info;
    }
    // cppLasso / cppOMP (cpp_library/cppspams.h:45-51, :106-113)
    Matrix<double>* cpath = nullptr;
    SpMatrix<double>* cl = cppLasso(X, D, &cpath, false, 10, 0.15);
    write_csc("cpp_lasso", *cl);
    delete cl;
    const int L8 = 8;
    double eps0 = 0, lam0 = 0;
    SpMatrix<double> co;
    cppOMP(X, D, co, &L8, &eps0, &lam0);
    write_csc("cpp_omp", co);
    // lasso2 with D given (decomp/decomp.h:126-129), L = 10
    SpMatrix<double> l2;
    lasso2(X, D, l2, 10, 0.15);
    write_csc("lasso2_d", l2);
  } catch (const char* msg) {
    std::fprintf(stderr, "dropin_run: %s\n", msg);
    return 5;
  }
  std::printf("dropin_run: ok\n");
  return 0;
}
