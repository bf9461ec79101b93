// Real-space coupling tables of unit spheres (host, FP64).
//
// Follows real_space_table of "CPU - Ewald test/ewald.py":604-931 (also
// Capacitance_Hetero/RealSpaceTable.m).  Every coupling is
//   sum over s in {+2, -2, 0} of  E_s(r) exp(-(r+s)^2 xi^2) + F_s(r) erfc((r+s) xi)
//   + [r < 2] R(r)
// with E_s, F_s = scale(xi) * poly(r) / r^k.  The closed-form coefficients are
// the reference's (they define the method); they are stored here as ascending
// polynomial coefficient lists and evaluated by Horner's rule.
#include <cmath>
#include <cstdint>
#include <vector>

namespace mpol {
namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

struct Rational {
  double scale;
  int k;                     // divide by r^k
  std::vector<double> c;     // ascending powers of r
  double eval(double r) const {
    double acc = 0.0;
    for (size_t i = c.size(); i-- > 0;) acc = acc * r + c[i];
    return scale * acc / std::pow(r, k);
  }
};

struct Coupling {
  Rational ep, em, e0, fp, fm, f0;
  int reg;  // which overlap regularisation
};

double regularisation(int which, double r) {
  switch (which) {
    case 0: return -1.0 / (4 * kPi * r) + (4.0 - r) / (16 * kPi);
    case 1: return -1.0 / (4 * kPi * r * r) + r / (8 * kPi) * (1.0 - 3.0 * r / 8.0);
    case 2: return -1.0 / (4 * kPi * r * r * r) + (1.0 - 9.0 * r / 16.0 + r * r * r / 32.0) / (4 * kPi);
    case 3: return 1.0 / (2 * kPi * r * r * r) + (1.0 - 9.0 * r / 8.0 + r * r * r / 8.0) / (4 * kPi);
    case 4: return 3.0 / (4 * kPi * r * r * r * r) - 3.0 / (64 * kPi) * (3.0 - r * r / 2.0);
    default: return -9.0 / (4 * kPi * r * r * r * r) - 9.0 / (64 * kPi) * (1.0 - r * r / 2.0);
  }
}

std::vector<Coupling> couplings(double x) {
  const double p32 = std::pow(kPi, 1.5);
  const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2;
  const double a = 3.0 + 4.0 * x2 - 32.0 * x4;
  const double b = 3.0 - 12.0 * x2 + 32.0 * x4;
  std::vector<Coupling> out(6);
  // potential / charge
  out[0] = {{1.0 / (32 * p32 * x), 1, {-2.0, -1.0}},
            {1.0 / (32 * p32 * x), 1, {2.0, -1.0}},
            {1.0 / (16 * p32 * x), 0, {1.0}},
            {1.0 / (64 * kPi * x2), 1, {8 * x2 + 1, 8 * x2, 2 * x2}},
            {1.0 / (64 * kPi * x2), 1, {8 * x2 + 1, -8 * x2, 2 * x2}},
            {1.0 / (32 * kPi * x2), 1, {-1.0, 0.0, -2 * x2}},
            0};
  // potential / dipole (= field / charge)
  out[1] = {{1.0 / (256 * p32 * x3), 2, {2 * (1 - 8 * x2), -3 + 8 * x2, -4 * x2, -6 * x2}},
            {1.0 / (256 * p32 * x3), 2, {-2 * (1 - 8 * x2), -3 + 8 * x2, 4 * x2, -6 * x2}},
            {3.0 / (128 * p32 * x3), 1, {1.0, 0.0, 2 * x2}},
            {1.0 / (512 * kPi * x4), 2, {-3 + 64 * x4, 0.0, 12 * x2, 32 * x4, 12 * x4}},
            {1.0 / (512 * kPi * x4), 2, {-3 + 64 * x4, 0.0, 12 * x2, -32 * x4, 12 * x4}},
            {3.0 / (256 * kPi * x4), 2, {1.0, 0.0, -4 * x2, 0.0, -4 * x4}},
            1};
  // field / dipole, I - rr
  out[2] = {{1.0 / (1024 * p32 * x5), 3, {2 * a, b, -8 * x2 * (3 + 2 * x2), 8 * x2 * (2 - 7 * x2), -8 * x4, 4 * x4}},
            {1.0 / (1024 * p32 * x5), 3, {-2 * a, b, 8 * x2 * (3 + 2 * x2), 8 * x2 * (2 - 7 * x2), 8 * x4, 4 * x4}},
            {1.0 / (512 * p32 * x5), 2, {-3 + 36 * x2, 0.0, -8 * x2 * (2 - 9 * x2), 0.0, -4 * x4}},
            {1.0 / (2048 * kPi * x6), 3,
             {3 - 36 * x2 + 256 * x6, 0.0, -18 * x2 * (1 - 8 * x2), 256 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -8 * x6}},
            {1.0 / (2048 * kPi * x6), 3,
             {3 - 36 * x2 + 256 * x6, 0.0, -18 * x2 * (1 - 8 * x2), -256 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -8 * x6}},
            {1.0 / (1024 * kPi * x6), 3,
             {-3 + 36 * x2, 0.0, 18 * x2 * (1 - 8 * x2), 0.0, 36 * x4 * (1 - 4 * x2), 0.0, 8 * x6}},
            2};
  // field / dipole, rr
  out[3] = {{1.0 / (512 * p32 * x5), 3, {-2 * a, -b, -4 * x2 * (3 - 4 * x2), 2 * x2 * (7 - 20 * x2), -16 * x4, 8 * x4}},
            {1.0 / (512 * p32 * x5), 3, {2 * a, -b, 4 * x2 * (3 - 4 * x2), 2 * x2 * (7 - 20 * x2), 16 * x4, 8 * x4}},
            {1.0 / (256 * p32 * x5), 2, {3 - 36 * x2, 0.0, -2 * x2 * (7 - 36 * x2), 0.0, -8 * x4}},
            {1.0 / (1024 * kPi * x6), 3,
             {-3 + 36 * x2 - 256 * x6, 0.0, 0.0, 128 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -16 * x6}},
            {1.0 / (1024 * kPi * x6), 3,
             {-3 + 36 * x2 - 256 * x6, 0.0, 0.0, -128 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -16 * x6}},
            {1.0 / (512 * kPi * x6), 3, {3 - 36 * x2, 0.0, 0.0, 0.0, 36 * x4 * (1 - 4 * x2), 0.0, 16 * x6}},
            3};
  // force coefficient of -(mi.mj) r and companions
  out[4] = {{3.0 / (1024 * p32 * x5), 4, {-2 * a, -b, 16 * x4, 4 * x2 * (1 - 2 * x2), -8 * x4, 4 * x4}},
            {3.0 / (1024 * p32 * x5), 4, {2 * a, -b, -16 * x4, 4 * x2 * (1 - 2 * x2), 8 * x4, 4 * x4}},
            {3.0 / (512 * p32 * x5), 3, {3 - 36 * x2, 0.0, -4 * x2 * (1 - 6 * x2), 0.0, -4 * x4}},
            {3.0 / (2048 * kPi * x6), 4,
             {-3 + 36 * x2 - 256 * x6, 0.0, 6 * x2 * (1 - 8 * x2), 0.0, -12 * x4 * (1 - 4 * x2), 0.0, -8 * x6}},
            {3.0 / (2048 * kPi * x6), 4,
             {-3 + 36 * x2 - 256 * x6, 0.0, 6 * x2 * (1 - 8 * x2), 0.0, -12 * x4 * (1 - 4 * x2), 0.0, -8 * x6}},
            {3.0 / (1024 * kPi * x6), 4,
             {3 - 36 * x2, 0.0, -6 * x2 * (1 - 8 * x2), 0.0, 12 * x4 * (1 - 4 * x2), 0.0, 8 * x6}},
            4};
  // force coefficient of -(mi.r)(mj.r) r
  out[5] = {{9.0 / (1024 * p32 * x5), 4, {2 * a, b, 8 * x2 * (1 - 2 * x2), 8 * x4, -8 * x4, 4 * x4}},
            {9.0 / (1024 * p32 * x5), 4, {-2 * a, b, -8 * x2 * (1 - 2 * x2), 8 * x4, 8 * x4, 4 * x4}},
            {9.0 / (512 * p32 * x5), 3, {-3 + 36 * x2, 0.0, 8 * x4, 0.0, -4 * x4}},
            {9.0 / (2048 * kPi * x6), 4,
             {3 - 36 * x2 + 256 * x6, 0.0, -2 * x2 * (1 - 8 * x2), 0.0, -4 * x4 * (1 - 4 * x2), 0.0, -8 * x6}},
            {9.0 / (2048 * kPi * x6), 4,
             {3 - 36 * x2 + 256 * x6, 0.0, -2 * x2 * (1 - 8 * x2), 0.0, -4 * x4 * (1 - 4 * x2), 0.0, -8 * x6}},
            {9.0 / (1024 * kPi * x6), 4,
             {-3 + 36 * x2, 0.0, 2 * x2 * (1 - 8 * x2), 0.0, 4 * x4 * (1 - 4 * x2), 0.0, 8 * x6}},
            5};
  return out;
}

}  // namespace

// out: 6 x (n+1); entry 0 of each row is the self term.
void real_space_tables(double xi, const double* r, int64_t n, double* out) {
  const std::vector<Coupling> cp = couplings(xi);
  const double e4 = std::exp(-4.0 * xi * xi);
  const double tail = std::erfc(2.0 * xi) / (4.0 * kPi);
  const double p32 = std::pow(kPi, 1.5);
  const double self_pc = (1.0 - e4) / (8.0 * p32 * xi) + tail;
  const double self_fd = (-1.0 + 6.0 * xi * xi + (1.0 - 2.0 * xi * xi) * e4) / (16.0 * p32 * xi * xi * xi) + tail;
  const double selfs[6] = {self_pc, 0.0, self_fd, self_fd, 0.0, 0.0};
  for (int t = 0; t < 6; ++t) out[t * (n + 1)] = selfs[t];
  for (int64_t i = 0; i < n; ++i) {
    const double rv = r[i];
    const double ep = std::exp(-(rv + 2.0) * (rv + 2.0) * xi * xi);
    const double em = std::exp(-(rv - 2.0) * (rv - 2.0) * xi * xi);
    const double e0 = std::exp(-rv * rv * xi * xi);
    const double fp = std::erfc((rv + 2.0) * xi);
    const double fm = std::erfc((rv - 2.0) * xi);
    const double f0 = std::erfc(rv * xi);
    for (int t = 0; t < 6; ++t) {
      const Coupling& c = cp[t];
      double v = c.ep.eval(rv) * ep + c.em.eval(rv) * em + c.e0.eval(rv) * e0 + c.fp.eval(rv) * fp +
                 c.fm.eval(rv) * fm + c.f0.eval(rv) * f0;
      if (rv < 2.0) v += regularisation(c.reg, rv);
      out[t * (n + 1) + 1 + i] = v;
    }
  }
}

}  // namespace mpol
