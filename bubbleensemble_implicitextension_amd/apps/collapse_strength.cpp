// Collapse-strength scan on the MI355X solver, written against the drop-in ProblemSolver
// header exactly like the reference's driver (CollapseStrength.cu:68-204): a PA x f grid of
// systems of UPS coupled Keller-Miksis bubbles, NTRN transient periods, an accessory reset,
// NCOLL recording periods, and one output line per (system, collapse) with the bubbles'
// maxima and following minima of R/R_E.
//
// Build (done by bubbleensemble_implicitextension_amd/build.py):
//   g++ -O2 -std=c++17 -Iinclude collapse_strength.cpp -L<package dir> -lbex -o collapse_strength
// Run:
//   collapse_strength <radii file> <coupling file> <output file> [PAMIN PAMAX FRMIN FRMAX
//                     [HEAVY MEDIUM]]
// With HEAVY and MEDIUM the same scan runs on the split solver (SplitProblemSolver: the HEAVY
// systems with the largest amplitude / frequency and the next MEDIUM one per wave, the others
// two per wave, all on the plain BiCG start); its result file equals the plain one.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <fstream>
#include <iostream>
#include <vector>

#include "CoupledSystems_PerBlock_Interface.hpp"

#ifndef PARES
#define PARES 8
#endif
#ifndef FRRES
#define FRRES 8
#endif
#ifndef NTRN
#define NTRN 4
#endif
#ifndef NCOLL
#define NCOLL 32
#endif
#ifndef CSTR
#define CSTR 1.0
#endif
#ifndef DEVICE
#define DEVICE 0
#endif

const int NS = PARES * FRRES;
const int UPS = 32;
const int UD = 2;
const int TPB = 32;
const int SPB = 1;
const int NC = 1;
const int CBW = 0;
const int CCI = 0;
const int NUP = 21;
const int NSP = 0;
const int NGP = 0;
const int NiGP = 0;
const int NUA = 2 * NCOLL;
const int NiUA = 1;
const int NSA = 0;
const int NiSA = 0;
const int NE = 1;
const int NDO = 0;

using Solver = ProblemSolver<NS, UPS, UD, TPB, SPB, NC, CBW, CCI, NUP, NSP, NGP, NiGP, NUA, NiUA,
                             NSA, NiSA, NE, NDO, RKCK45, double>;
using Split = SplitProblemSolver<NS, UPS, UD, TPB, SPB, NC, CBW, CCI, NUP, NSP, NGP, NiGP, NUA,
                                 NiUA, NSA, NiSA, NE, NDO, RKCK45, double>;

static std::vector<double> grid(double lo, double hi, int n, bool logarithmic) {
  std::vector<double> v(n, lo);
  if (n == 1) return v;
  for (int i = 0; i < n; i++) {
    double t = (double)i / (n - 1);
    v[i] = logarithmic ? std::pow(10.0, std::log10(lo) + t * (std::log10(hi) - std::log10(lo)))
                       : lo + t * (hi - lo);
  }
  v[n - 1] = hi;
  return v;
}

// Per-unit Keller-Miksis constants C0..C20 (CollapseStrength.cu:321-357).
template <class S>
static void fill_system(S& s, int sid, double pa_bar, double f_khz,
                        const std::vector<double>& radii_um) {
  const double PI = 3.14159265358979323846;
  const double Pv = 3.166775638952003e+03, Rho = 9.970639504998557e+02;
  const double ST = 0.071977583160056, Vis = 8.902125058209557e-04;
  const double CL = 1.497251785455527e+03, Pinf = 1.0e5, gamma = 1.4;
  s.SetHost(sid, TimeDomain, 0, 0.0);
  s.SetHost(sid, TimeDomain, 1, 1.0);
  s.SetHost(sid, CouplingStrength, 0, 1.0);
  for (int uid = 0; uid < UPS; uid++) {
    const double RE = radii_um[uid] / 1e6;
    const double w1 = 2.0 * PI * (f_khz * 1000);
    const double L = 2.0 * PI / RE / w1;
    const double PA1 = pa_bar * 1e5;
    const double C[NUP] = {(2.0 * ST / RE + Pinf - Pv) * L * L / Rho,
                           (1.0 - 3.0 * gamma) * (2 * ST / RE + Pinf - Pv) * L / CL / Rho,
                           (Pinf - Pv) * L * L / Rho,
                           (2.0 * ST / RE / Rho) * L * L,
                           (4.0 * Vis / Rho / (RE * RE)) * (2.0 * PI / w1),
                           PA1 * L * L / Rho,
                           0.0,
                           (RE * w1 * PA1 / Rho / CL) * L * L,
                           0.0,
                           RE * w1 / (2.0 * PI) / CL,
                           3.0 * gamma,
                           0.0,
                           0.0,
                           2.0 * PI / w1,
                           RE,
                           pa_bar,
                           f_khz,
                           0.0,
                           0.0,
                           0.0,
                           radii_um[uid]};
    s.SetHost(sid, uid, ActualState, 0, 1.0);
    s.SetHost(sid, uid, ActualState, 1, 0.0);
    for (int k = 0; k < NUP; k++) s.SetHost(sid, uid, UnitParameters, k, C[k]);
    s.SetHost(sid, uid, IntegerUnitAccessories, 0, -1);
  }
}

template <class S>
static int run(S& s, const std::vector<double>& radii, const std::vector<double>& M,
               const std::vector<double>& amp, const std::vector<double>& freq,
               const char* output);

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: " << argv[0]
              << " radii.txt coupling.txt output.txt [PAMIN PAMAX FRMIN FRMAX [HEAVY MEDIUM]]"
              << std::endl;
    return 2;
  }
  double pamin = argc > 4 ? std::atof(argv[4]) : 0.5, pamax = argc > 5 ? std::atof(argv[5]) : 2.0;
  double frmin = argc > 6 ? std::atof(argv[6]) : 20.0, frmax = argc > 7 ? std::atof(argv[7]) : 200.0;

  std::vector<double> radii(UPS);
  std::vector<double> M(UPS * UPS);
  {
    std::ifstream fr(argv[1]), fm(argv[2]);
    for (int i = 0; i < UPS; i++) fr >> radii[i];
    for (int i = 0; i < UPS * UPS; i++) fm >> M[i];  // row-major text, row = target bubble
    if (!fr || !fm) {
      std::cerr << "could not read the radii or coupling file" << std::endl;
      return 2;
    }
  }
  std::vector<double> amp = grid(pamin, pamax, PARES, false);
  std::vector<double> freq = grid(frmin, frmax, FRRES, false);

  ListCUDADevices();  // CollapseStrength.cu:80-81
  PrintPropertiesOfSpecificDevice(DEVICE);

  auto pa_of = [&](int sid) { return amp[sid % PARES]; };
  auto f_of = [&](int sid) { return freq[sid / PARES]; };
  if (argc > 9) {
    // Split by forcing amplitude over frequency, largest first.
    const int heavy = std::atoi(argv[8]), medium = std::atoi(argv[9]);
    if (heavy < 0 || medium < 0 || heavy + medium > NS) {
      std::cerr << "HEAVY and MEDIUM must be >= 0 with HEAVY + MEDIUM <= " << NS << std::endl;
      return 2;
    }
    std::vector<int> order(NS);
    for (int i = 0; i < NS; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return pa_of(a) / f_of(a) > pa_of(b) / f_of(b);
    });
    std::vector<int> h(order.begin(), order.begin() + heavy);
    std::vector<int> m(order.begin() + heavy, order.begin() + heavy + medium);
    Split s(DEVICE, h, m);
    std::cout << "split solver: " << s.Parts() << " parts" << std::endl;
    return run(s, radii, M, amp, freq, argv[3]);
  }
  Solver s(DEVICE);
  return run(s, radii, M, amp, freq, argv[3]);
}

template <class S>
static int run(S& s, const std::vector<double>& radii, const std::vector<double>& M,
               const std::vector<double>& amp, const std::vector<double>& freq,
               const char* output) {
  s.SolverOption(SharedCouplingMatrices, 0);
  s.SolverOption(InitialTimeStep, 1e-3);
  s.SolverOption(RelativeTolerance, 0, 1e-10);
  s.SolverOption(RelativeTolerance, 1, 1e-10);
  s.SolverOption(AbsoluteTolerance, 0, 1e-10);
  s.SolverOption(AbsoluteTolerance, 1, 1e-10);
  s.SolverOption(MinimumTimeStep, 1e-13);
  s.SolverOption(EventDirection, 0, -1);
  s.SolverOption(EventTolerance, 0, 1e-4);

  int sid = 0;
  for (double f : freq)  // frequency outer, amplitude inner (CollapseStrength.cu:306-308)
    for (double pa : amp) fill_system(s, sid++, pa, f, radii);
  for (int r = 0; r < UPS; r++)
    for (int c = 0; c < UPS; c++) s.SetHost(0, CouplingMatrix, r, c, CSTR * M[r * UPS + c]);
  s.SetHost(CouplingIndex, 0, 1);

  auto t0 = std::chrono::steady_clock::now();
  s.SynchroniseFromHostToDevice(All);
  for (int i = 0; i < NTRN; i++) {
    s.Solve();
    s.InsertSynchronisationPoint();
    s.SynchroniseSolver();
  }
  for (int sys = 0; sys < NS; sys++)
    for (int uid = 0; uid < UPS; uid++) {
      for (int k = 0; k < 2 * NCOLL; k++) s.SetHost(sys, uid, UnitAccessories, k, 0.0);
      s.SetHost(sys, uid, IntegerUnitAccessories, 0, -1);
    }
  s.SynchroniseFromHostToDevice(UnitAccessories);
  s.SynchroniseFromHostToDevice(IntegerUnitAccessories);
  for (int i = 0; i < NCOLL; i++) {
    s.Solve();
    s.InsertSynchronisationPoint();
    s.SynchroniseSolver();
  }
  s.SynchroniseFromDeviceToHost(All);
  s.InsertSynchronisationPoint();
  s.SynchroniseSolver();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  std::ofstream out(output);
  out.flags(std::ios::scientific);
  for (int sys = 0; sys < NS; sys++)
    for (int c = 0; c < NCOLL; c++) {
      out.width(6);
      out << sys << ',';
      out.width(17);
      out.precision(10);
      out << s.template GetHost<double>(sys, 0, UnitParameters, 15) << ',';
      out.width(17);
      out << s.template GetHost<double>(sys, 0, UnitParameters, 16) << ',';
      for (int uid = 0; uid < UPS; uid++) {
        out.width(10);
        out.precision(3);
        out << s.template GetHost<double>(sys, uid, UnitAccessories, c) << ',';
      }
      for (int uid = 0; uid < UPS; uid++) {
        out.width(10);
        out.precision(3);
        out << s.template GetHost<double>(sys, uid, UnitAccessories, c + NCOLL) << ',';
      }
      out << '\n';
    }
  std::cout << "systems " << NS << ", periods " << NTRN + NCOLL << ", wall " << secs << " s"
            << std::endl;
  return 0;
}
