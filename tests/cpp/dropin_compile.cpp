// Compiles the spams.h-compatible templates the way the reference's C++
// harness (cpp_library/test_decomp.cpp) uses them.
#include "spams_amd.h"

void use_lasso() {
  Matrix<double> X(100, 10), D(100, 200);
  Matrix<double>* path;
  SpMatrix<double>* spa = cppLasso(X, D, &path, false, 10, 0.15);
  delete spa;
  Matrix<float> Xf(64, 5), Df(64, 256);
  SpMatrix<float>* s2 = _lassoD<float>(&Xf, &Df, reinterpret_cast<Matrix<float>**>(&path), false,
                                       -1, 0.15f, 0.f, PENALTY, false, false, -1, -1, false, false);
  delete s2;
  // return_reg_path and ols, and the Gram-given form
  SpMatrix<float>* s3 = _lassoD<float>(&Xf, &Df, reinterpret_cast<Matrix<float>**>(&path), true,
                                       -1, 0.15f, 0.f, PENALTY, false, true, -1, -1, false, false);
  delete s3;
  delete reinterpret_cast<Matrix<float>*>(path);
  Matrix<float> Q(256, 256), q(256, 5);
  SpMatrix<float>* s4 = _lassoQq<float>(&Xf, &Q, &q, reinterpret_cast<Matrix<float>**>(&path),
                                        false, -1, 0.15f, 0.f, PENALTY, false, false, -1, -1,
                                        false, false);
  delete s4;
  Matrix<float> Wf(256, 5);
  SpMatrix<float>* s5 = _lassoWeighted<float>(&Xf, &Df, &Wf, -1, 0.15f, PENALTY, false, -1, false);
  delete s5;
}

void use_omp() {
  Matrix<double> X(64, 10), D(64, 200);
  const int L = 10;
  double eps = 1.0, lambda = 0.;
  SpMatrix<double> spa;
  cppOMP(X, D, spa, &L, &eps, &lambda);
  Vector<int> vL(1);
  vL[0] = 5;
  Vector<double> ve(1), vl(1);
  Matrix<double>* path;
  SpMatrix<double>* s = _omp(&X, &D, &path, false, true, &vL, false, &ve, false, &vl, -1);
  delete s;
  SpMatrix<double>* s2 = _omp(&X, &D, &path, true, true, &vL, false, &ve, false, &vl, -1);
  delete s2;
  delete path;
  Matrix<bool> B(64, 10);
  SpMatrix<double>* s3 = _ompMask(&X, &D, &B, &path, false, true, &vL, false, &ve, false, &vl, -1);
  delete s3;
  SpMatrix<double>* s4 = _lassoMask(&X, &D, &B, 10, 0.15, 0.0, PENALTY, false, -1, false);
  delete s4;
}

void use_fista() {
  Matrix<double> Y(32, 4), D(32, 64), W0(64, 4), W(64, 4);
  Vector<int> groups;
  Vector<double> iw;
  char reg[] = "l1", loss[] = "square", logn[] = "";
  Matrix<double>* info = _fistaFlat(&Y, &D, &W0, &W, &groups, -1, 100, 0.1, false, 1.5, 0.05,
                                    1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 1e-6, 10, 1000, false, false,
                                    false, false, false, reg, loss, false, false, false, false,
                                    false, false, logn, false, &iw, 1, true, false, 0);
  delete info;
}

void use_train_dl() {
  Matrix<float> X(64, 1000), D0;
  Matrix<float> mA, mB;
  Matrix<float>* A = nullptr;
  Matrix<float>* B = nullptr;
  Vector<int>* it = nullptr;
  Vector<float> eta;
  SpMatrix<bool> g, gv;
  Vector<int> own, nown;
  char regul[] = "none";
  char logname[] = "";
  Matrix<float>* D = _alltrainDL<float>(&X, false, &A, &B, &it, true, &mA, &mB, 0, &D0, &eta, &g,
                                        &gv, &own, &nown, -1, 1e-6f, true, false, 512, 128, 0.15,
                                        1e-9, 0., 10, 1e-5, PENALTY, regul, false, false, false,
                                        L2, false, true, false, 0., 0., 1.0f, 1, false, 0, false,
                                        false, logname);
  delete D;
  delete A;
  delete B;
  delete it;
}

void use_decomp_level() {
  Matrix<float> X(64, 20), D(64, 256), G(256, 256), DtX(256, 20);
  SpMatrix<float> a1, a2, a3, a4;
  lasso(X, D, a1, 64, 0.15f);                         // decomp/decomp.h:111-116
  lasso(X, D, a2, 64, 0.15f, 0.f, PENALTY, false, true);
  lasso2(X, G, DtX, a3, 64, 0.15f);                   // decomp/decomp.h:131-135
  lasso(X, G, DtX, a4, 64, 0.15f, L1COEFFS, true, false);
}
