// Real-space coupling tables of unit spheres (host, FP64).
//
// Follows real_space_table of "CPU - Ewald test/ewald.py":604-931 (also
// Capacitance_Hetero/RealSpaceTable.m), and the eight couplings that
// Capacitance_Quadrupole/RealSpaceTable.m adds for quadrupoles (:66-85,
// :173-356, self term :372).  Every coupling is
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

double regularisation(int which, double r, double x = 0.0) {
  const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r;
  switch (which) {
    case 6: {  // potential / quadrupole (the reference's form, which depends on xi)
      const double x2 = x * x;
      return 15.0 / (1024 * kPi * std::pow(x, 6) * r3) * (12 * x2 * x2 * r4 - 2 * x2 * (3 + 8 * x2) * r2 + 3 + 12 * x2);
    }
    case 7: return -15.0 / (4 * kPi * r4) + 15.0 * r2 / (64 * kPi) * (1.0 - 3.0 * r2 / 16.0);
    case 8: return 3.0 / (4 * kPi * r4) - 3.0 * r / (8 * kPi) * (1.0 - 5.0 * r / 8.0 + 5.0 * r3 / 128.0);
    case 9: return 5.0 * r / (8 * kPi) * (1.0 - 3.0 * r / 4.0 + r3 / 16.0);
    case 10: return 75.0 * r * (4.0 - r2) * (4.0 - r2) / (1024 * kPi);
    case 11: return -3.0 / (2 * kPi) * (1.0 / r5 - 1.0 + 25.0 * r / 32.0 - 25.0 * r3 / 256.0 + 3.0 * r5 / 512.0);
    case 12: return -15.0 * (r2 - 4.0) * (r2 - 4.0) * (r2 - 4.0) * (3 * r4 + 6 * r2 + 8) / (2048 * kPi * r5);
    case 13: return -15.0 * (3584.0 - 80 * r4 * r2 - 30 * r4 * r4 + 9 * r5 * r5) / (2048 * kPi * r5);
    default: break;
  }
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

// The eight couplings of the dipole + quadrupole method, in the order
// pot_quad, field_quad_1, field_quad_2, field_quad_3, grad_quad_1..4.
std::vector<Coupling> quad_couplings(double x) {
  const double p32 = std::pow(kPi, 1.5);
  const double x2 = x * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2, x7 = x6 * x, x8 = x4 * x4, x9 = x8 * x,
               x10 = x8 * x2;
  std::vector<Coupling> out(8);
  // potential / quadrupole
  {
    const double a = 9 - 4 * x2 + 32 * x4, b = 15 - 12 * x2 + 32 * x4, c = 16 * x2 * (2 - x2), d = 2 * x2 * (10 - 4 * x2);
    out[0] = {{3.0 / (1024 * p32 * x5), 3, {2 * a, -b, -c, d, 24 * x4, 20 * x4}},
              {3.0 / (1024 * p32 * x5), 3, {-2 * a, -b, c, d, -24 * x4, 20 * x4}},
              {15.0 / (512 * p32 * x5), 2, {-3 * (1 + 4 * x2), 0.0, 4 * x2 * (1 + 2 * x2), 0.0, 4 * x4}},
              {3.0 / (2048 * kPi * x6), 3,
               {15 + 60 * x2 + 256 * x6, 0.0, -10 * x2 * (3 + 8 * x2), 0.0, 20 * x4 * (3 + 4 * x2), 128 * x6, 40 * x6}},
              {3.0 / (2048 * kPi * x6), 3,
               {15 + 60 * x2 + 256 * x6, 0.0, -10 * x2 * (3 + 8 * x2), 0.0, 20 * x4 * (3 + 4 * x2), -128 * x6, 40 * x6}},
              {15.0 / (1024 * kPi * x6), 3,
               {-3 * (1 + 4 * x2), 0.0, -2 * x2 * (3 + 8 * x2), 0.0, -4 * x4 * (3 + 4 * x2), 0.0, -8 * x6}},
              6};
  }
  // field / quadrupole (= field gradient / dipole): rrr.S
  {
    const double a = 45 + 24 * x2 - 64 * x4 + 512 * x6, b = 45 - 120 * x2 + 192 * x4 - 512 * x6;
    const double c = 4 * x2 * (3 - 8 * x2) * (3 - 8 * x2), d = 2 * x2 * (21 - 80 * x2 + 64 * x4);
    const double e = 8 * x4 * (3 - 8 * x2), f = -4 * x4 * (9 - 8 * x2);
    const std::vector<double> fpm = {-45 + 480 * x2 + 4096 * x8, 0.0, 72 * x2 * (1 - 8 * x2), 0.0,
                                     -24 * x4 * (3 - 16 * x2), 0.0, 32 * x6 * (3 - 8 * x2), 0.0, 48 * x8};
    out[1] = {{15.0 / (16384 * p32 * x7), 4, {-2 * a, -b, -c, d, e, f, 48 * x6, -24 * x6}},
              {15.0 / (16384 * p32 * x7), 4, {2 * a, -b, c, d, -e, f, -48 * x6, -24 * x6}},
              {15.0 / (8192 * p32 * x7), 3,
               {45 - 480 * x2, 0.0, -2 * x2 * (21 - 128 * x2), 0.0, 4 * x4 * (9 - 32 * x2), 0.0, 24 * x6}},
              {15.0 / (32768 * kPi * x8), 4, fpm},
              {15.0 / (32768 * kPi * x8), 4, fpm},
              {15.0 / (16384 * kPi * x8), 4,
               {45 - 480 * x2, 0.0, -72 * x2 * (1 - 8 * x2), 0.0, 24 * x4 * (3 - 16 * x2), 0.0, -32 * x6 * (3 - 8 * x2),
                0.0, -48 * x8}},
              7};
  }
  // Ir.S + S r + r S
  {
    const double a = 45 + 24 * x2 - 64 * x4 + 512 * x6, b = 45 - 120 * x2 + 192 * x4 - 512 * x6;
    const double c = 4 * x2 * (15 + 48 * x2 - 64 * x4), d = -2 * x2 * (45 - 80 * x2 + 64 * x4);
    const double e = 8 * x4 * (45 + 8 * x2), f = -20 * x4 * (11 - 24 * x2);
    out[2] = {{3.0 / (16384 * p32 * x7), 4, {2 * a, b, -c, d, e, f, 80 * x6, -40 * x6}},
              {3.0 / (16384 * p32 * x7), 4, {-2 * a, b, c, d, -e, f, -80 * x6, -40 * x6}},
              {15.0 / (8192 * p32 * x7), 3,
               {-9 + 96 * x2, 0.0, 2 * x2 * (9 - 64 * x2), 0.0, 4 * x4 * (11 - 32 * x2), 0.0, 8 * x6}},
              {3.0 / (32768 * kPi * x8), 4,
               {45 - 480 * x2 - 4096 * x8, 0.0, -120 * x2 * (1 - 8 * x2), 0.0, 120 * x4 * (3 - 16 * x2), -2048 * x8,
                160 * x6 * (3 - 8 * x2), 0.0, 80 * x8}},
              {3.0 / (32768 * kPi * x8), 4,
               {45 - 480 * x2 - 4096 * x8, 0.0, -120 * x2 * (1 - 8 * x2), 0.0, 120 * x4 * (3 - 16 * x2), 2048 * x8,
                160 * x6 * (3 - 8 * x2), 0.0, 80 * x8}},
              {15.0 / (16384 * kPi * x8), 4,
               {-9 + 96 * x2, 0.0, 24 * x2 * (1 - 8 * x2), 0.0, -24 * x4 * (3 - 16 * x2), 0.0, -32 * x6 * (3 - 8 * x2),
                0.0, -16 * x8}},
              8};
  }
  // Ir.S
  out[3] = {{5.0 / (1024 * p32 * x5), 2, {6.0, 3.0, -24 * x2, 16 * x2 * (1 - 2 * x2), -8 * x4, 4 * x4}},
            {5.0 / (1024 * p32 * x5), 2, {-6.0, 3.0, 24 * x2, 16 * x2 * (1 - 2 * x2), 8 * x4, 4 * x4}},
            {5.0 / (512 * p32 * x5), 1, {-3 + 24 * x2, 0.0, -16 * x2 * (1 - 3 * x2), 0.0, -4 * x4}},
            {5.0 / (2048 * kPi * x6), 2,
             {3 - 24 * x2, 0.0, -6 * x2 * (3 - 16 * x2), 128 * x6, -12 * x4 * (3 - 8 * x2), 0.0, -8 * x6}},
            {5.0 / (2048 * kPi * x6), 2,
             {3 - 24 * x2, 0.0, -6 * x2 * (3 - 16 * x2), -128 * x6, -12 * x4 * (3 - 8 * x2), 0.0, -8 * x6}},
            {5.0 / (1024 * kPi * x6), 2,
             {-3 + 24 * x2, 0.0, 6 * x2 * (3 - 16 * x2), 0.0, 12 * x4 * (3 - 8 * x2), 0.0, 8 * x6}},
            9};
  // field gradient / quadrupole: Irr:Q
  {
    const std::vector<double> g = {9 - 48 * x2 + 192 * x4, 0.0, -(24 * x2 - 96 * x4 + 256 * x6), 0.0,
                                   72 * x4 - 192 * x6 + 256 * x8, 0.0, 96 * x6 - 128 * x8, 0.0, 16 * x8};
    out[4] = {{75.0 / (16384 * p32 * x7), 3,
               {-(18 - 48 * x2), -(9 + 24 * x2), 12 * x2 - 64 * x4, 18 * x2 + 32 * x4, -(72 * x4 - 64 * x6),
                44 * x4 - 32 * x6, -16 * x6, 8 * x6}},
              {75.0 / (16384 * p32 * x7), 3,
               {18 - 48 * x2, -(9 + 24 * x2), -(12 * x2 - 64 * x4), 18 * x2 + 32 * x4, 72 * x4 - 64 * x6,
                44 * x4 - 32 * x6, 16 * x6, 8 * x6}},
              {-75.0 / (8192 * p32 * x7), 2,
               {-9 + 48 * x2 - 192 * x4, 0.0, 18 * x2 - 64 * x4 + 128 * x6, 0.0, 44 * x4 - 64 * x6, 0.0, 8 * x6}},
              {-75.0 / (32768 * kPi * x8), 3, g},
              {-75.0 / (32768 * kPi * x8), 3, g},
              {75.0 / (16384 * kPi * x8), 3, g},
              10};
  }
  // Q
  {
    const double a = 270 - 1080 * x2 - 192 * x4 + 512 * x6 - 4096 * x8;
    const double b = 135 + 180 * x2 + 480 * x4 - 768 * x6 + 2048 * x8;
    const double c = 360 * x2 - 1680 * x4 - 768 * x6 + 1024 * x8;
    const double d = -(360 * x2 + 360 * x4 + 640 * x6 - 512 * x8);
    const double e = 2160 * x4 - 4320 * x6 - 256 * x8;
    const double f = -(1536 * x4 - 2576 * x6 + 3968 * x8);
    const double g = 1056 * x6 - 1216 * x8;
    const double h = -(576 * x6 - 608 * x8);
    std::vector<double> fc = {405 - 2700 * x2 + 14400 * x4, 0.0, -(1350 * x2 - 7200 * x4 + 28800 * x6), 0.0,
                              5400 * x4 - 21600 * x6 + 57600 * x8, 0.0, 10800 * x6 - 28800 * x8 + 38400 * x10, 0.0,
                              3600 * x8 - 4800 * x10, 0.0, 288 * x10};
    std::vector<double> fp = fc, fm = fc;
    fp[0] += 49152 * x10;
    fm[0] += 49152 * x10;
    fp[5] = 49152 * x10;
    fm[5] = -49152 * x10;
    out[5] = {{3.0 / (32768 * p32 * x9), 5, {a, b, -c, d, e, f, g, h, 96 * x8, -48 * x8}},
              {3.0 / (32768 * p32 * x9), 5, {-a, b, c, d, -e, f, -g, h, -96 * x8, -48 * x8}},
              {1.0 / (16384 * p32 * x9), 4,
               {-405 + 2700 * x2 - 14400 * x4, 0.0, 1080 * x2 - 5400 * x4 + 19200 * x6, 0.0,
                4608 * x4 - 13200 * x6 + 19200 * x8, 0.0, 1728 * x6 - 2400 * x8, 0.0, 144 * x8}},
              {1.0 / (65536 * kPi * x10), 5, fp},
              {1.0 / (65536 * kPi * x10), 5, fm},
              {-1.0 / (32768 * kPi * x10), 5, fc},
              11};
  }
  // Irr:Q + 2 Q.rr + 2 rr.Q
  {
    const double a = -(-270 + 1080 * x2 + 192 * x4 - 512 * x6 + 4096 * x8);
    const double b = 135 + 180 * x2 + 480 * x4 - 768 * x6 + 2048 * x8;
    const double c = 720 * x4 + 768 * x6 - 1024 * x8;
    const double d = -(180 * x2 - 120 * x4 + 640 * x6 - 512 * x8);
    const double e = -(480 * x6 + 256 * x8);
    const double f = 216 * x4 + 144 * x6 + 128 * x8;
    const double g = -576 * (x6 - x8);
    const double h = 336 * x6 - 288 * x8;
    const std::vector<double> fpm = {135 - 900 * x2 + 4800 * x4 + 16384 * x10, 0.0,
                                     -(270 * x2 - 1440 * x4 + 5760 * x6), 0.0, 360 * x4 - 1440 * x6 + 3840 * x8, 0.0,
                                     -(720 * x6 - 1920 * x8 + 2560 * x10), 0.0, -720 * x8 + 960 * x10, 0.0,
                                     -96 * x10};
    out[6] = {{-15.0 / (65536 * p32 * x9), 5, {a, b, c, d, e, f, g, h, -96 * x8, 48 * x8}},
              {-15.0 / (65536 * p32 * x9), 5, {-a, b, -c, d, -e, f, -g, h, 96 * x8, 48 * x8}},
              {15.0 / (32768 * p32 * x9), 4,
               {135 - 900 * x2 + 4800 * x4, 0.0, -(180 * x2 - 840 * x4 + 2560 * x6), 0.0,
                216 * x4 - 720 * x6 + 1280 * x8, 0.0, 336 * x6 - 480 * x8, 0.0, 48 * x8}},
              {-15.0 / (131072 * kPi * x10), 5, fpm},
              {-15.0 / (131072 * kPi * x10), 5, fpm},
              {-15.0 / (65536 * kPi * x10), 5,
               {-135 + 900 * x2 - 4800 * x4, 0.0, 270 * x2 - 1440 * x4 + 5760 * x6, 0.0,
                -(360 * x4 - 1440 * x6 + 3840 * x8), 0.0, 720 * x6 - 1920 * x8 + 2560 * x10, 0.0,
                720 * x8 - 960 * x10, 0.0, 96 * x10}},
              12};
  }
  // rrrr:Q
  {
    const double a = -(1890 - 7560 * x2 - 1344 * x4 + 3584 * x6 - 28672 * x8);
    const double b = -(945 + 1260 * x2 + 3360 * x4 - 5376 * x6 + 14336 * x8);
    const double c = -(1080 * x2 + 2160 * x4 + 5376 * x6 - 7168 * x8);
    const double d = 720 * x2 - 2280 * x4 + 4480 * x6 - 3584 * x8;
    const double e = 720 * x4 - 480 * x6 + 1792 * x8;
    const double f = -(432 * x4 - 912 * x6 + 896 * x8);
    const double g = -(288 * x6 + 192 * x8);
    const double h = 288 * x6 + 96 * x8;
    std::vector<double> fc = {945 - 6300 * x2 + 33600 * x4, 0.0, -(1350 * x2 - 7200 * x4 + 28800 * x6), 0.0,
                              1080 * x4 - 4320 * x6 + 11520 * x8, 0.0, -(720 * x6 - 1920 * x8 + 2560 * x10), 0.0,
                              720 * x8 - 960 * x10, 0.0, 288 * x10};
    std::vector<double> fpm = fc;
    fpm[0] += 114688 * x10;
    out[7] = {{-15.0 / (65536 * p32 * x9), 5, {a, b, c, d, e, f, g, h, -288 * x8, 144 * x8}},
              {-15.0 / (65536 * p32 * x9), 5, {-a, b, -c, d, -e, f, -g, h, 288 * x8, 144 * x8}},
              {15.0 / (32768 * p32 * x9), 4,
               {-945 + 6300 * x2 - 33600 * x4, 0.0, 720 * x2 - 3000 * x4 + 6400 * x6, 0.0,
                -(432 * x4 - 1200 * x6 + 1280 * x8), 0.0, 288 * x6 - 480 * x8, 0.0, 144 * x8}},
              {15.0 / (131072 * kPi * x10), 5, fpm},
              {15.0 / (131072 * kPi * x10), 5, fpm},
              {-15.0 / (65536 * kPi * x10), 5, fc},
              13};
  }
  return out;
}

}  // namespace

// out: 8 x (n+1) in the order pot_quad, field_quad_1..3, grad_quad_1..4;
// entry 0 of each row is the self term (only grad_quad_2 has one).
// regularize = false leaves out the [r < 2] overlap terms, as the field-map
// tables do (Compute_Field_Quadrupole/RealSpaceTable.m).
void real_space_tables_quad(double xi, const double* r, int64_t n, double* out, bool regularize) {
  const std::vector<Coupling> cp = quad_couplings(xi);
  const double sp = std::sqrt(kPi), x2 = xi * xi, x5 = x2 * x2 * xi;
  const double self_gq2 = 1.0 / (2 * kPi) *
                          (3 * (3 - 10 * x2 + 20 * x2 * x2) / (8 * sp * x5) -
                           3 * (3 + 2 * x2 + 4 * x2 * x2) * std::exp(-4 * x2) / (8 * sp * x5) + 3 * std::erfc(2 * xi));
  for (int t = 0; t < 8; ++t) out[t * (n + 1)] = (t == 5) ? self_gq2 : 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double rv = r[i];
    const double ep = std::exp(-(rv + 2.0) * (rv + 2.0) * xi * xi);
    const double em = std::exp(-(rv - 2.0) * (rv - 2.0) * xi * xi);
    const double e0 = std::exp(-rv * rv * xi * xi);
    const double fp = std::erfc((rv + 2.0) * xi);
    const double fm = std::erfc((rv - 2.0) * xi);
    const double f0 = std::erfc(rv * xi);
    for (int t = 0; t < 8; ++t) {
      const Coupling& c = cp[t];
      double v = c.ep.eval(rv) * ep + c.em.eval(rv) * em + c.e0.eval(rv) * e0 + c.fp.eval(rv) * fp +
                 c.fm.eval(rv) * fm + c.f0.eval(rv) * f0;
      if (regularize && rv < 2.0) v += regularisation(c.reg, rv, xi);
      out[t * (n + 1) + 1 + i] = v;
    }
  }
}

// out: 6 x (n+1); entry 0 of each row is the self term.  regularize = false
// leaves out the [r < 2] overlap terms, as the field-map tables do
// (Compute_Field/RealSpaceTable.m: the couplings of a point to a sphere).
void real_space_tables(double xi, const double* r, int64_t n, double* out, bool regularize) {
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
      if (regularize && rv < 2.0) v += regularisation(c.reg, rv);
      out[t * (n + 1) + 1 + i] = v;
    }
  }
}

}  // namespace mpol
