// Calls the library's host code through the C ABI under AddressSanitizer /
// UndefinedBehaviorSanitizer, where no device context is needed:
//   * version, device count, context creation refused for a bad device or a
//     null output pointer, destroy / last_error of a null context;
//   * every entry point that takes a context returns SVH_ERR_ARG for a null
//     one (abi_null_ctx.inc, generated from include/svh.h);
//   * svh_matern_general: argument checks, and its values against the closed
//     forms at nu = 1/2, 3/2, 5/2, 7/2, at both sides of the series / continued
//     fraction switch, and g = M'(r)/r against a central difference of M.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "svh.h"

static int failures = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

static double closed_form(double nu, double r) {
  if (nu == 0.5) return std::exp(-r);
  const double c = std::sqrt(2.0 * nu) * r;
  if (nu == 1.5) return (1.0 + c) * std::exp(-c);
  if (nu == 2.5) return (1.0 + c + c * c / 3.0) * std::exp(-c);
  return (1.0 + c + 2.0 * c * c / 5.0 + c * c * c / 15.0) * std::exp(-c);  // nu = 7/2
}

static void matern_checks() {
  double r0 = 0.5, M0 = 0.0, g0 = 0.0;
  CHECK(svh_matern_general(0.0, &r0, 1, &M0, &g0) == SVH_ERR_ARG);
  CHECK(svh_matern_general(-1.0, &r0, 1, &M0, &g0) == SVH_ERR_ARG);
  CHECK(svh_matern_general(std::nan(""), &r0, 1, &M0, &g0) == SVH_ERR_ARG);
  CHECK(svh_matern_general(1.0, &r0, -1, &M0, &g0) == SVH_ERR_ARG);
  CHECK(svh_matern_general(1.0, nullptr, 1, &M0, &g0) == SVH_ERR_ARG);
  CHECK(svh_matern_general(1.0, &r0, 1, nullptr, &g0) == SVH_ERR_ARG);
  CHECK(svh_matern_general(1.0, nullptr, 0, nullptr, nullptr) == SVH_OK);

  std::vector<double> r;
  for (int e = -9; e <= 1; ++e)
    for (double f : {1.0, 2.0, 5.0}) r.push_back(f * std::pow(10.0, e));
  r.push_back(0.0);
  r.push_back(60.0);
  std::vector<double> M(r.size()), g(r.size());
  for (double nu : {0.5, 1.5, 2.5, 3.5}) {
    // both sides of the switch at sqrt(2 nu) r = 2
    std::vector<double> rr = r;
    const double rs = 2.0 / std::sqrt(2.0 * nu);
    rr.push_back(rs * (1.0 - 1e-12));
    rr.push_back(rs * (1.0 + 1e-12));
    M.resize(rr.size());
    g.resize(rr.size());
    CHECK(svh_matern_general(nu, rr.data(), (int64_t)rr.size(), M.data(), g.data()) == SVH_OK);
    for (size_t i = 0; i < rr.size(); ++i) {
      const double ref = rr[i] < 1e-7 ? 1.0 : closed_form(nu, rr[i]);
      CHECK(std::fabs(M[i] - ref) <= 1e-12 * std::fabs(ref) + 1e-300);
    }
    const size_t k = rr.size();
    CHECK(std::fabs(M[k - 1] - M[k - 2]) <= 1e-11);
  }
  for (double nu : {0.2, 0.7, 1.0, 1.9, 3.3, 10.0}) {
    std::vector<double> rr = {0.05, 0.3, 0.9, 1.7, 3.0};
    std::vector<double> rp(rr), rm(rr), Mp(rr.size()), Mm(rr.size()), Mx(rr.size()), gx(rr.size()),
        gp(rr.size()), gm(rr.size());
    const double h = 1e-5;
    for (size_t i = 0; i < rr.size(); ++i) {
      rp[i] = rr[i] + h;
      rm[i] = rr[i] - h;
    }
    CHECK(svh_matern_general(nu, rr.data(), (int64_t)rr.size(), Mx.data(), gx.data()) == SVH_OK);
    CHECK(svh_matern_general(nu, rp.data(), (int64_t)rr.size(), Mp.data(), gp.data()) == SVH_OK);
    CHECK(svh_matern_general(nu, rm.data(), (int64_t)rr.size(), Mm.data(), gm.data()) == SVH_OK);
    for (size_t i = 0; i < rr.size(); ++i) {
      CHECK(Mx[i] > 0.0 && Mx[i] < 1.0 && gx[i] < 0.0);
      const double fd = (Mp[i] - Mm[i]) / (2.0 * h) / rr[i];
      CHECK(std::fabs(gx[i] - fd) <= 1e-6 * (1.0 + std::fabs(fd)));
    }
  }
}

int main() {
  const char* v = svh_version();
  CHECK(v && std::strstr(v, "gfx950"));
  const int nd = svh_device_count();
  CHECK(nd >= 0);
  CHECK(svh_create(0, nullptr) == SVH_ERR_ARG);
  svh_ctx* ctx = nullptr;
  CHECK(svh_create(-1, &ctx) != SVH_OK && ctx == nullptr);
  CHECK(svh_create(nd, &ctx) != SVH_OK && ctx == nullptr);
  svh_destroy(nullptr);
  CHECK(std::strcmp(svh_last_error(nullptr), "") == 0);

#include "abi_null_ctx.inc"

  matern_checks();
  std::printf("abi_checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
