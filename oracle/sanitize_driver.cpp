// AddressSanitizer / UndefinedBehaviorSanitizer driver for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: built together with bex_oracle.cpp under the sanitizers by
// `make -C oracle sanitize`, run by the tests. It reads one problem written by the tests
// (dimensions, options, buffers in the reference's SoA layout, the file format below), solves it
// once in each arithmetic mode on two threads and writes the outputs, so that the tests can
// hold them to the regular build bit for bit. A sanitizer finding aborts the process with a
// nonzero status.
//
// Input file (native endianness), in order:
//   int32  dims[14]      NS N UD NUP NSP NGP NiGP NUA NiUA NSA NiSA NE NDO NC
//   double opt_d[8]      InitialTimeStep MaximumTimeStep MinimumTimeStep TimeStepGrowLimit
//                        TimeStepShrinkLimit DenseOutputMinimumTimeStep BiCGRhoTolerance
//                        BiCGResidualTolerance
//   int32  opt_i[5]      ActiveSystems DenseOutputSaveFrequency BiCGMaxIterations Algorithm
//                        DenseOutputInterpolation
//   double RelativeTolerance[UD], AbsoluteTolerance[UD], EventTolerance[NE]
//   int32  EventDirection[NE]
//   then the buffers of bex_oracle_buffers in declaration order, each at its full size.
// Output file: for mode 0 then mode 1, ActualState, ActualTime, UnitAccessories,
// IntegerUnitAccessories, DenseOutputTimeInstances, DenseOutputStates, DenseOutputIndex, stats.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "bex_oracle.h"

namespace {

struct Reader {
  FILE* f;
  bool ok = true;
  template <class T>
  std::vector<T> take(size_t n) {
    std::vector<T> v(n > 0 ? n : 1);
    if (n > 0 && std::fread(v.data(), sizeof(T), n, f) != n) ok = false;
    return v;
  }
};

template <class T>
void put(FILE* f, const std::vector<T>& v, size_t n) {
  if (n > 0) std::fwrite(v.data(), sizeof(T), n, f);
}

size_t at_least_one(size_t n) { return n > 0 ? n : 1; }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s problem.bin outputs.bin\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 3;
  Reader r{in};
  const std::vector<int32_t> d = r.take<int32_t>(14);
  bex_oracle_dims dims{d[0], d[1], d[2], d[3], d[4], d[5], d[6],
                       d[7], d[8], d[9], d[10], d[11], d[12], d[13]};
  const std::vector<double> od = r.take<double>(8);
  const std::vector<int32_t> oi = r.take<int32_t>(5);
  const std::vector<double> rtol = r.take<double>(dims.UD);
  const std::vector<double> atol = r.take<double>(dims.UD);
  const std::vector<double> etol = r.take<double>(at_least_one(dims.NE));
  const std::vector<int32_t> edir = r.take<int32_t>(at_least_one(dims.NE));

  const size_t NS = dims.NS, TLT = (size_t)dims.NS * dims.N;
  const size_t n_td = 2 * NS, n_x = dims.UD * TLT, n_t = NS, n_up = dims.NUP * TLT;
  const size_t n_sp = at_least_one(dims.NSP * NS), n_gp = at_least_one(dims.NGP);
  const size_t n_igp = at_least_one(dims.NiGP), n_ua = dims.NUA * TLT, n_iua = dims.NiUA * TLT;
  const size_t n_sa = at_least_one(dims.NSA * NS), n_isa = at_least_one(dims.NiSA * NS);
  const size_t n_m = (size_t)dims.NC * dims.N * dims.N;
  const size_t n_dt = at_least_one((size_t)dims.NDO * NS);
  const size_t n_ds = at_least_one((size_t)dims.NDO * dims.UD * TLT);

  const auto td = r.take<double>(n_td);
  const auto x = r.take<double>(n_x);
  const auto t = r.take<double>(n_t);
  const auto up = r.take<double>(n_up);
  const auto sp = r.take<double>(n_sp);
  const auto gp = r.take<double>(n_gp);
  const auto igp = r.take<int32_t>(n_igp);
  const auto ua = r.take<double>(n_ua);
  const auto iua = r.take<int32_t>(n_iua);
  const auto sa = r.take<double>(n_sa);
  const auto isa = r.take<int32_t>(n_isa);
  const auto m = r.take<double>(n_m);
  const auto dt = r.take<double>(n_dt);
  const auto ds = r.take<double>(n_ds);
  const auto di = r.take<int32_t>(NS);
  std::fclose(in);
  if (!r.ok) {
    std::fprintf(stderr, "short problem file\n");
    return 4;
  }

  bex_oracle_options o{};
  o.InitialTimeStep = od[0];
  o.MaximumTimeStep = od[1];
  o.MinimumTimeStep = od[2];
  o.TimeStepGrowLimit = od[3];
  o.TimeStepShrinkLimit = od[4];
  o.DenseOutputMinimumTimeStep = od[5];
  o.BiCGRhoTolerance = od[6];
  o.BiCGResidualTolerance = od[7];
  o.ActiveSystems = oi[0];
  o.DenseOutputSaveFrequency = oi[1];
  o.BiCGMaxIterations = oi[2];
  o.Algorithm = oi[3];
  o.DenseOutputInterpolation = oi[4];
  o.RelativeTolerance = rtol.data();
  o.AbsoluteTolerance = atol.data();
  o.EventTolerance = etol.data();
  o.EventDirection = edir.data();

  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 5;
  for (int mode = 0; mode < 2; ++mode) {
    // Fresh copies per mode: a solve writes its state, accessories and dense output back.
    auto td_m = td, x_m = x, t_m = t, up_m = up, sp_m = sp, gp_m = gp, ua_m = ua, sa_m = sa,
         m_m = m, dt_m = dt, ds_m = ds;
    auto igp_m = igp, iua_m = iua, isa_m = isa, di_m = di;
    std::vector<long long> stats(NS * BEX_NSTAT, 0);
    bex_oracle_buffers b{td_m.data(), x_m.data(),  t_m.data(),   up_m.data(), sp_m.data(),
                         gp_m.data(), igp_m.data(), ua_m.data(), iua_m.data(), sa_m.data(),
                         isa_m.data(), m_m.data(),  dt_m.data(), ds_m.data(), di_m.data()};
    const int rc = bex_oracle_solve(&dims, &o, &b, stats.data(), mode, 2);
    if (rc != 0) {
      std::fprintf(stderr, "bex_oracle_solve returned %d in mode %d\n", rc, mode);
      return 6;
    }
    put(out, x_m, n_x);
    put(out, t_m, n_t);
    put(out, ua_m, n_ua);
    put(out, iua_m, n_iua);
    put(out, dt_m, n_dt);
    put(out, ds_m, n_ds);
    put(out, di_m, NS);
    put(out, stats, stats.size());
  }
  std::fclose(out);
  return 0;
}
